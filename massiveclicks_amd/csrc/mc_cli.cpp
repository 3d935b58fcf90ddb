// mc_cli.cpp -- command-line EM driver on one MI355X.
//
// Takes the reference's command line (main.cpp:222-236) and prints its
// report (main.cpp:326-441 and parallel_em.cu:603-608, 697-740) for a single
// worker, with the work done by libmclicks.so:
//   parse + split + sort + qd indices   Dataset::parse_dataset
//   n EM iterations                      em_parallel (parallel_em.cu:461-609)
//   held-out LL and perplexity           evaluation.cpp:25-66
//   parameter CSV files (-o)             Communicate::output_parameters
// Multi-node and multi-GPU flags are accepted with their single-worker
// meaning; CPU-only execution (-e 1) is not implemented (the engine is GPU-only).

#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "mclicks.hpp"

namespace {

struct Config {
    std::string dataset_path = "YandexRelPredChallenge100k.txt";   // main.cpp:201
    std::string output_path;
    std::string checkpoint_in, checkpoint_out;
    int thread_count = (int)std::thread::hardware_concurrency();
    int max_gpus = 1;
    int iterations = 50;
    int max_sessions = 40000;
    int model_type = 0;
    int partitioning_type = 0;
    int job_id = 0;
    int exec_mode = 0;
    int reduce_mode = MC_REDUCE_SEGMENTED;
    double test_share = 0.2;
    bool help = false;
    bool autotune = false;
    std::string fit = "em";
    double lr = 0.05;
    int batches = 16;
};

void show_help() {
    std::cout <<
        "Usage: mclicks [options]\n"
        "  -r, --raw-path <file>      raw click log (default YandexRelPredChallenge100k.txt)\n"
        "  -s, --max-sessions <n>     sessions to read, -1 = all (default 40000)\n"
        "  -i, --itr <n>              EM iterations (default 50)\n"
        "  -m, --model-type <id>      0 PBM, 1 CCM, 2 DBN, 3 UBM (default 0)\n"
        "  -t, --test-share <f>       share of sessions held out (default 0.2)\n"
        "  -o, --output <prefix>      write <prefix>_<parameter>.csv files\n"
        "  -g, --n-gpus <n>           GPUs to use (must be 1)\n"
        "  -n, --n-threads <n>        host threads (reported)\n"
        "  -p, --partition-type <id>  multi-GPU partitioning (no effect with one GPU)\n"
        "  -j, --job-id <id>          job id (reported)\n"
        "  -e, --exec-mode <id>       0 GPU-only (the only mode)\n"
        "      --reduce-mode <id>     0 deterministic segmented (default), 1 float atomics\n"
        "      --autotune <0|1>       measure and pick the EM launch options first\n"
        "      --fit <em|adam>        plain EM (default) or Adam on the log-likelihood gradient;\n"
        "                             with adam, -i counts epochs (every minibatch once)\n"
        "      --lr <f>               Adam learning rate (default 0.05)\n"
        "      --batches <n>          Adam minibatches of whole query groups (default 16)\n"
        "      --resume <file>        load a parameter checkpoint before training\n"
        "      --checkpoint <file>    save a parameter checkpoint after training\n"
        "  -h, --help                 this message\n";
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
    const auto t_total = std::chrono::steady_clock::now();
    Config config;
    struct Option {
        std::string long_form, short_form;
        std::function<void(const char*)> handler;
    };
    const std::vector<Option> options = {
        {"--n-threads", "-n", [&](const char* v) { config.thread_count = std::stoi(v); }},
        {"--n-gpus", "-g", [&](const char* v) { config.max_gpus = std::stoi(v); }},
        {"--raw-path", "-r", [&](const char* v) { config.dataset_path = v; }},
        {"--output", "-o", [&](const char* v) { config.output_path = v; }},
        {"--itr", "-i", [&](const char* v) { config.iterations = std::stoi(v); }},
        {"--max-sessions", "-s", [&](const char* v) { config.max_sessions = std::stoi(v); }},
        {"--model-type", "-m", [&](const char* v) { config.model_type = std::stoi(v); }},
        {"--partition-type", "-p", [&](const char* v) { config.partitioning_type = std::stoi(v); }},
        {"--test-share", "-t", [&](const char* v) { config.test_share = std::stod(v); }},
        {"--job-id", "-j", [&](const char* v) { config.job_id = std::stoi(v); }},
        {"--exec-mode", "-e", [&](const char* v) { config.exec_mode = std::stoi(v); }},
        {"--reduce-mode", "", [&](const char* v) { config.reduce_mode = std::stoi(v); }},
        {"--autotune", "", [&](const char* v) { config.autotune = std::stoi(v) != 0; }},
        {"--fit", "", [&](const char* v) { config.fit = v; }},
        {"--lr", "", [&](const char* v) { config.lr = std::stod(v); }},
        {"--batches", "", [&](const char* v) { config.batches = std::stoi(v); }},
        {"--resume", "", [&](const char* v) { config.checkpoint_in = v; }},
        {"--checkpoint", "", [&](const char* v) { config.checkpoint_out = v; }},
        {"--help", "-h", [&](const char*) { config.help = true; }},
    };
    for (int i = 1; i < argc; i += 2) {
        const std::string name = argv[i];
        const char* value = i + 1 < argc ? argv[i + 1] : "";
        bool handled = false;
        try {
            for (const Option& o : options) {
                if (name == o.long_form || (!o.short_form.empty() && name == o.short_form)) {
                    o.handler(value);
                    handled = true;
                    break;
                }
            }
            if (!handled) {
                std::cerr << "Error: Did not recognize flag '" << name << "'." << std::endl;
                config.help = true;
            }
        } catch (std::exception&) {
            std::cerr << "Error: Invalid argument '" << value << "' for flag '" << name << "'." << std::endl;
            config.help = true;
        }
    }
    if (config.help) {
        show_help();
        return 0;
    }
    if (config.thread_count < 1) {                      // main.cpp:42-51
        std::cerr << "Error: Invalid number of threads: " << config.thread_count << std::endl;
        return 1;
    }
    if (config.exec_mode != 0) {
        std::cerr << "Error: only GPU-only execution (-e 0) is supported." << std::endl;
        return 1;
    }
    if (config.max_gpus != 1) {
        std::cerr << "Error: training runs on one GPU (-g 1); query-sharded multi-GPU training is not implemented."
                  << std::endl;
        return 1;
    }
    if (config.fit != "em" && config.fit != "adam") {
        std::cerr << "Error: --fit must be em or adam." << std::endl;
        return 1;
    }
    if (config.model_type < 0 || config.model_type > 3) {
        std::cerr << "Error: Invalid model type: " << config.model_type << std::endl;
        return 1;
    }
    if (config.partitioning_type < 0 || config.partitioning_type > 5) {   // main.cpp:77-85
        std::cerr << "Error: Invalid partitioning scheme: " << config.partitioning_type << std::endl;
        return 1;
    }

    const std::map<int, std::string> models{{0, "PBM"}, {1, "CCM"}, {2, "DBN"}, {3, "UBM"}};
    const std::map<int, std::string> partitionings{{0, "Round-Robin"},
                                                   {1, "Maximum Utilization"},
                                                   {2, "Proportional Maximum Utilization"},
                                                   {3, "Newest Architecture First"},
                                                   {4, "Relative CUDA Cores"},
                                                   {5, "Relative Peak Performance"}};
    char host[HOST_NAME_MAX + 1] = {0};
    gethostname(host, HOST_NAME_MAX);

    std::cout << "Job ID: " << config.job_id << "\nNumber of machines: 1\nNumber of workers: 1"
              << "\nNumber of threads: " << config.thread_count << "\nExecution mode: GPU-only"
              << "\nRaw data path: " << config.dataset_path << "\nNumber of EM iterations: " << config.iterations
              << "\nShare of data used for testing: " << config.test_share * 100 << "%"
              << "\nMax number of sessions: " << config.max_sessions
              << "\nPartitioning type: " << partitionings.at(config.partitioning_type)
              << "\nModel type: " << models.at(config.model_type) << std::endl
              << std::endl;

    hipDeviceProp_t prop;
    std::memset(&prop, 0, sizeof prop);
    size_t free_b = 0, total_b = 0;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        std::cerr << "Error: No GPU devices found for GPU-only execution." << std::endl;
        return 1;
    }
    std::cout << std::left << std::setw(std::max<int>((int)std::strlen(host), 4) + 2) << "Node" << std::setw(24)
              << "Device" << std::right << std::setw(8) << "Arch" << std::setw(14) << "Memory (MB)" << std::endl;
    std::cout << std::left << std::setw(std::max<int>((int)std::strlen(host), 4) + 2) << host << std::setw(24)
              << prop.name << std::right << std::setw(8) << prop.gcnArchName << std::setw(14) << std::fixed
              << std::setprecision(2) << free_b / 1e6 << std::endl
              << std::setprecision(8) << std::endl;
    std::cout.unsetf(std::ios::floatfield);

    try {
        // ---- preprocessing (main.cpp:373-408) -------------------------------
        const auto t_pre = std::chrono::steady_clock::now();
        std::cout << "Parsing dataset..." << std::endl;
        mclicks::Dataset ds = mclicks::Dataset::parse(config.dataset_path, config.max_sessions,
                                                      (float)config.test_share);
        std::cout << std::left << std::setw(std::max<int>((int)std::strlen(host), 4)) << "Node" << std::right
                  << std::setw(12) << "Train" << std::setw(12) << "Test" << std::setw(12) << "QD Pairs" << std::endl
                  << std::left << std::setw(std::max<int>((int)std::strlen(host), 4)) << host << std::right
                  << std::setw(12) << ds.n_train() << std::setw(12) << ds.n_test() << std::setw(12) << ds.n_qd()
                  << std::endl;
        const double pre_s = seconds_since(t_pre);

        // ---- EM (parallel_em.cu:349-609) ------------------------------------
        const auto t_em = std::chrono::steady_clock::now();
        int64_t opts[MC_N_OPTIONS] = {config.reduce_mode, 8192, 1280, 1, 256, 0, 1, 1, 1, 1, 1};
        if (config.autotune) {
            const mclicks::Sessions tr = ds.train();
            double ms = 0.0;
            int trials = 0;
            mclicks::check(mc_autotune(config.model_type, 0, tr.size(), tr.query.data(), tr.pidx.data(),
                                       tr.clicks.data(), nullptr, ds.n_qd(), 5, opts, &ms, &trials));
            std::cout << "\nAutotune (" << trials << " candidates): reduce-mode " << opts[MC_OPT_REDUCE_MODE]
                      << ", item-sessions " << opts[MC_OPT_ITEM_SESSIONS] << ", window " << opts[MC_OPT_ITEM_WINDOW]
                      << ", persistent " << opts[MC_OPT_PERSISTENT] << ", block " << opts[MC_OPT_BLOCK] << ": " << ms
                      << " ms/iteration" << std::endl;
        }
        if (config.fit == "adam") opts[MC_OPT_BATCHES] = std::max(1, config.batches);
        mclicks::ClickModel cm(config.model_type, 0);
        for (int o = 0; o < MC_N_OPTIONS; o++) cm.set_option(o, opts[o]);
        const auto t_h2d = std::chrono::steady_clock::now();
        cm.init_parameters(ds);
        const double h2d_s = seconds_since(t_h2d);
        if (!config.checkpoint_in.empty()) cm.load(config.checkpoint_in);
        const int itr_len = (int)std::to_string(config.iterations).length() + 1;
        double sum_itr = 0, sum_comp = 0, sum_upd = 0;
        if (config.fit == "adam") {
            std::cout << "\nStarting " << config.iterations << " Adam epochs (" << opts[MC_OPT_BATCHES]
                      << " minibatches, lr " << config.lr << ")..." << std::endl;
            cm.grad_init();
            cm.adam_configure((float)config.lr);
        } else {
            std::cout << "\nStarting " << config.iterations << " EM parameter estimation iterations..." << std::endl;
        }
        for (int itr = 0; itr < config.iterations && config.fit == "adam"; itr++) {
            double ms = 0.0;
            mclicks::check(mc_adam_steps_timed(cm.handle(), (int)opts[MC_OPT_BATCHES], &ms));
            sum_itr += ms / 1e3;
            sum_comp += ms / 1e3;
            std::cout << "Epoch: " << std::left << std::setw(itr_len) << itr << " Epoch-time: " << std::setw(11)
                      << ms / 1e3 << std::right << std::endl;
        }
        for (int itr = 0; itr < config.iterations && config.fit == "em"; itr++) {
            double ms_total = 0, ms_e = 0, ms_m = 0;
            mclicks::check(mc_em_iterations_timed(cm.handle(), 1, &ms_total, &ms_e, &ms_m));
            sum_itr += ms_total / 1e3;
            sum_comp += ms_e / 1e3;
            sum_upd += ms_m / 1e3;
            std::cout << "Itr: " << std::left << std::setw(itr_len) << itr << " Itr-time: " << std::setw(11)
                      << ms_total / 1e3 << " Itr-EM_COMP: " << std::setw(11) << ms_e / 1e3
                      << " Itr-EM_UPDATE: " << std::setw(11) << ms_m / 1e3 << " Itr-Sync: " << std::setw(12) << 0.0
                      << std::right << std::endl;
        }

        // ---- evaluation (parallel_em.cu:633-659) ----------------------------
        const auto t_eval = std::chrono::steady_clock::now();
        mclicks::Evaluation ev = cm.evaluate(ds.test());
        const double eval_s = seconds_since(t_eval);

        if (!config.output_path.empty()) {                 // parallel_em.cu:677-687
            std::cout << "\nWriting output to file..." << std::endl;
            cm.output_parameters(ds, config.output_path);
        }
        if (!config.checkpoint_out.empty()) cm.save(config.checkpoint_out);

        std::cout << "\nTotal Log likelihood is: " << ev.log_likelihood() << std::endl;
        for (int r = 0; r < MC_MAX_SERP; r++)
            std::cout << "Perplexity at rank " << r << " is: " << ev.perplexity_at_rank(r) << std::endl;
        std::cout << "Perplexity is: " << ev.perplexity() << std::endl;
        const int n = std::max(config.iterations, 1);
        std::cout << "\nHost to Device dataset transfer time: " << h2d_s
                  << "\nAverage time per iteration: " << sum_itr / n
                  << "\nAverage time per computation in each iteration: " << sum_comp / n
                  << "\nAverage time per update in each iteration: " << sum_upd / n
                  << "\nAverage time per synchronization in each iteration: " << 0.0
                  << "\nTotal time of training: " << sum_itr << "\nEvaluation time: " << eval_s << std::endl;

        const double em_s = seconds_since(t_em);
        const double total_s = seconds_since(t_total);
        auto line = [](const char* name, double v, double tot) {
            std::cout << std::left << std::setw(27) << name << std::fixed << std::setprecision(7) << v
                      << " seconds, " << std::right << std::setw(3) << std::setprecision(0)
                      << std::round(v / tot * 100) << " %" << std::endl;
        };
        std::cout << std::endl;
        line("Pre-processing time: ", pre_s, pre_s + em_s);
        line("Parameter estimation time: ", em_s, pre_s + em_s);
        line("Total elapsed time: ", total_s, total_s);
        std::cout << std::endl;
    } catch (mclicks::Error& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
