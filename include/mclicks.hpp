// mclicks.hpp -- C++ host interface over the C ABI in mclicks.h.
//
// Shaped like the reference's host classes so that a C++ driver written
// against them maps call for call:
//   mclicks::Param         Param (click_models/param.cuh:18-39)
//   mclicks::Dataset       Dataset::parse_dataset (data/dataset.cpp:421-455)
//                          for one worker: train set, test set, n_qd
//   mclicks::ClickModel    ClickModel_Hst (click_models/base.cuh:30-53),
//                          created by create_cm_host (base.cu:19-49)
//   mclicks::LogLikelihood / Perplexity  (click_models/evaluation.cpp:25-66)
// Header only; links against libmclicks.so.  Errors throw mclicks::Error
// (the reference aborts the process instead, utils/utils.cuh:27-42).
#ifndef MCLICKS_HPP
#define MCLICKS_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mclicks.h"

namespace mclicks {

// Parameter groups (utils/definitions.h:37-39).
constexpr int PUBLIC = 0;
constexpr int PRIVATE = 1;
constexpr int ALL = -1;

class Error : public std::runtime_error {
public:
    Error(mc_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    mc_status status;
};

inline void check(mc_status s) {
    if (s != MC_OK) throw Error(s, std::string("mclicks: ") + mc_last_error());
}

// (numerator, denominator) pair; value() as param.cu:75-87.
struct Param {
    float num = 1.f;
    float den = 2.f;
    float value() const {
        const float v = num / den;
        return v < 1.f - 0.000001f ? v : 1.f - 0.000001f;
    }
};

// Sessions of one set in the engine's host layout (session-major, 10 ranks).
struct Sessions {
    std::vector<int32_t> query;   // [n]
    std::vector<int32_t> docs;    // [n * 10]
    std::vector<uint8_t> clicks;  // [n * 10]
    std::vector<int32_t> pidx;    // [n * 10], -1 = pair unseen in training
    std::vector<uint8_t> length;  // [n] or empty (= 10 ranks)
    int64_t size() const { return (int64_t)query.size(); }
};

class Dataset {
public:
    // Dataset::parse_dataset for one worker (main.cpp -r, -s, -t).
    static Dataset parse(const std::string& path, int max_sessions, float test_share) {
        mc_dataset* h = nullptr;
        check(mc_dataset_parse(path.c_str(), max_sessions, test_share, &h));
        return Dataset(h);
    }
    static Dataset from_sessions(int64_t n, const int32_t* query, const int32_t* docs, const uint8_t* clicks,
                                 float test_share) {
        mc_dataset* h = nullptr;
        check(mc_dataset_from_sessions(n, query, docs, clicks, test_share, &h));
        return Dataset(h);
    }
    int64_t n_parsed() const { return sizes()[0]; }
    int64_t n_train() const { return sizes()[1]; }
    int64_t n_test() const { return sizes()[2]; }
    int64_t n_qd() const { return sizes()[3]; }
    Sessions train() const { return export_set(0, n_train()); }
    Sessions test() const { return export_set(1, n_test()); }
    const mc_dataset* handle() const { return h_.get(); }

private:
    explicit Dataset(mc_dataset* h) : h_(h, &mc_dataset_free) {}
    std::array<int64_t, 4> sizes() const {
        std::array<int64_t, 4> s{};
        check(mc_dataset_sizes(h_.get(), &s[0], &s[1], &s[2], &s[3]));
        return s;
    }
    Sessions export_set(int which, int64_t n) const {
        Sessions s;
        s.query.resize((size_t)n);
        s.docs.resize((size_t)n * MC_MAX_SERP);
        s.clicks.resize((size_t)n * MC_MAX_SERP);
        s.pidx.resize((size_t)n * MC_MAX_SERP);
        if (n) check(mc_dataset_export(h_.get(), which, s.query.data(), s.docs.data(), s.clicks.data(), s.pidx.data()));
        return s;
    }
    std::shared_ptr<mc_dataset> h_;
};

struct Evaluation {
    double llh_sum = 0.0;                  // sum over sessions of the per-session mean log-probability
    std::array<double, MC_MAX_SERP> ppl_log2_sum{};
    std::array<int64_t, MC_MAX_SERP> rank_count{};
    int64_t n_sessions = 0;
    // LogLikelihood::evaluate / parallel_em.cu:704.
    double log_likelihood() const { return n_sessions ? llh_sum / (double)n_sessions : NAN; }
    // parallel_em.cu:719-721.
    double perplexity_at_rank(int r) const {
        return rank_count[r] ? std::pow(2.0, -ppl_log2_sum[r] / (double)rank_count[r]) : NAN;
    }
    // parallel_em.cu:723: mean over ranks.
    double perplexity() const {
        double s = 0.0;
        for (int r = 0; r < MC_MAX_SERP; r++) s += perplexity_at_rank(r);
        return s / MC_MAX_SERP;
    }
};

class ClickModel {
public:
    // create_cm_host (base.cu:19-49): 0 PBM, 1 CCM, 2 DBN, 3 UBM.
    explicit ClickModel(int model_type, int device = 0) {
        mc_model* m = nullptr;
        check(mc_model_create(model_type, device, &m));
        m_.reset(m);
    }
    ClickModel(ClickModel&&) = default;
    ClickModel& operator=(ClickModel&&) = default;

    void set_option(int option, int64_t value) { check(mc_model_set_option(m_.get(), option, value)); }
    int model_type() const {
        int t = -1;
        check(mc_get_model_type(m_.get(), &t));
        return t;
    }

    // init_parameters (base.cuh:39): tables to (1,2), sessions uploaded.
    void init_parameters(const Sessions& train, int64_t n_qd) {
        check(mc_init_parameters(m_.get(), train.size(), train.query.data(), train.pidx.data(), train.clicks.data(),
                                 train.length.empty() ? nullptr : train.length.data(), n_qd));
    }
    void init_parameters(const Dataset& ds) { check(mc_init_from_dataset(m_.get(), ds.handle())); }

    // process_session (base.cuh:36): E-step over the training set.
    void process_session() { check(mc_process_session(m_.get())); }
    // reset_parameters + update_parameters (base.cuh:37,43): M-step.
    void update_parameters() { check(mc_update_parameters(m_.get())); }
    // The iteration loop of parallel_em.cu:461-609.
    void em(int iterations, bool use_graph = true) { check(mc_em_iterations(m_.get(), iterations, use_graph ? 1 : 0)); }
    void reset_parameters() { check(mc_reset_parameters(m_.get())); }

    // get_parameter_information (base.cuh:33): names of the public and the
    // private tables.
    std::pair<std::vector<std::string>, std::vector<std::string>> get_parameter_information() const {
        std::pair<std::vector<std::string>, std::vector<std::string>> h;
        for (int t = 0; t < n_tables(); t++) {
            const char* name = nullptr;
            int pub = 0;
            int64_t size = 0;
            check(mc_get_table_info(m_.get(), t, &name, &pub, &size));
            (pub ? h.first : h.second).push_back(name);
        }
        return h;
    }

    // get_parameters (base.cuh:40, e.g. pbm.cu:305-321): PUBLIC, PRIVATE or
    // ALL (public tables first, then private ones).
    void get_parameters(std::vector<std::vector<Param>>& dst, int parameter_type) const {
        dst.clear();
        for (int t : tables_of(parameter_type)) {
            int64_t size = table_size(t);
            std::vector<float> num((size_t)size), den((size_t)size);
            check(mc_get_parameters(m_.get(), t, num.data(), den.data()));
            std::vector<Param> v((size_t)size);
            for (int64_t i = 0; i < size; i++) v[(size_t)i] = Param{num[(size_t)i], den[(size_t)i]};
            dst.push_back(std::move(v));
        }
    }
    // set_parameters (base.cuh:40, e.g. pbm.cu:331-343), same grouping.
    void set_parameters(const std::vector<std::vector<Param>>& src, int parameter_type) {
        const std::vector<int> ts = tables_of(parameter_type);
        if (src.size() != ts.size()) throw Error(MC_ERR_INVALID_ARG, "mclicks: wrong number of tables");
        for (size_t k = 0; k < ts.size(); k++) {
            std::vector<float> num(src[k].size()), den(src[k].size());
            for (size_t i = 0; i < src[k].size(); i++) { num[i] = src[k][i].num; den[i] = src[k][i].den; }
            if ((int64_t)num.size() != table_size(ts[k])) throw Error(MC_ERR_INVALID_ARG, "mclicks: wrong table size");
            check(mc_set_parameters(m_.get(), ts[k], num.data(), den.data()));
        }
    }

    // get_log_conditional_click_probs / get_full_click_probs (base.cuh:51-52)
    // for a whole set at once: [n * 10] each, ranks past a session's length 0.
    void get_log_conditional_click_probs(const Sessions& s, std::vector<float>& out) {
        out.resize((size_t)s.size() * MC_MAX_SERP);
        run_eval(s, out.data(), nullptr, nullptr);
    }
    void get_full_click_probs(const Sessions& s, std::vector<float>& out) {
        out.resize((size_t)s.size() * MC_MAX_SERP);
        run_eval(s, nullptr, out.data(), nullptr);
    }
    // LogLikelihood + Perplexity (evaluation.cpp:25-66), totals in FP64.
    Evaluation evaluate(const Sessions& s) { return run_eval(s, nullptr, nullptr, nullptr); }

    // get_serp_probability (base.cuh:50) for a whole set: [n * 10].
    void get_serp_probability(const Sessions& s, std::vector<float>& out) {
        out.resize((size_t)s.size() * MC_MAX_SERP);
        if (!s.size()) return;
        check(mc_serp_probability(m_.get(), s.size(), s.pidx.data(), s.clicks.data(),
                                  s.length.empty() ? nullptr : s.length.data(), out.data()));
    }

    // The held-out set kept on the device (mc_set_test_set), then evaluated
    // at the current parameters as often as needed (mc_evaluate_test_set).
    void set_test_set(const Sessions& s) {
        check(mc_set_test_set(m_.get(), s.size(), s.pidx.data(), s.clicks.data(),
                              s.length.empty() ? nullptr : s.length.data()));
        n_test_ = s.size();
    }
    Evaluation evaluate_test_set() {
        Evaluation e;
        e.n_sessions = n_test_;
        check(mc_evaluate_test_set(m_.get(), nullptr, nullptr, nullptr, &e.llh_sum, e.ppl_log2_sum.data(),
                                   e.rank_count.data(), nullptr));
        return e;
    }

    size_t get_memory_usage() const {
        size_t b = 0;
        check(mc_get_memory_usage(m_.get(), &b));
        return b;
    }
    static size_t compute_memory_footprint(int model_type, int64_t n_sessions, int64_t n_qd) {
        size_t b = 0;
        check(mc_compute_memory_footprint(model_type, n_sessions, n_qd, &b));
        return b;
    }

    // Direct maximum-likelihood fitting beside EM (no counterpart in the
    // reference): logits from the current values, then Adam steps over the
    // minibatches set with set_option(MC_OPT_BATCHES, ...).
    void grad_init() { check(mc_grad_init(m_.get())); }
    // Training log-likelihood of minibatch `batch` (-1 = all) and its
    // gradient with respect to every table's logits (public table last).
    double loglik_gradient(std::vector<std::vector<float>>& grad, int batch = -1) {
        std::vector<float*> ptrs;
        grad.resize((size_t)n_tables());
        for (int t = 0; t < n_tables(); t++) {
            grad[(size_t)t].resize((size_t)table_size(t));
            ptrs.push_back(grad[(size_t)t].data());
        }
        double ll = 0.0;
        check(mc_loglik_gradient(m_.get(), batch, &ll, ptrs.data()));
        return ll;
    }
    void adam_configure(float lr = 0.05f, float beta1 = 0.9f, float beta2 = 0.999f, float eps = 1e-8f,
                        bool use_prior = true) {
        check(mc_adam_configure(m_.get(), lr, beta1, beta2, eps, use_prior ? 1 : 0));
    }
    void adam_steps(int n_steps, bool use_graph = true) { check(mc_adam_steps(m_.get(), n_steps, use_graph ? 1 : 0)); }

    // Communicate::output_parameters (communicator.cpp:415-516).
    void output_parameters(const Dataset& ds, const std::string& prefix) {
        check(mc_output_parameters(m_.get(), ds.handle(), prefix.c_str()));
    }
    void save(const std::string& path) { check(mc_save_parameters(m_.get(), path.c_str())); }
    void load(const std::string& path) { check(mc_load_parameters(m_.get(), path.c_str())); }

    mc_model* handle() const { return m_.get(); }

private:
    struct Deleter {
        void operator()(mc_model* m) const { mc_model_destroy(m); }
    };
    int n_tables() const {
        int n = 0;
        check(mc_get_n_tables(m_.get(), &n));
        return n;
    }
    int64_t table_size(int t) const {
        const char* name = nullptr;
        int pub = 0;
        int64_t size = 0;
        check(mc_get_table_info(m_.get(), t, &name, &pub, &size));
        return size;
    }
    std::vector<int> tables_of(int parameter_type) const {
        std::vector<int> pub, priv;
        for (int t = 0; t < n_tables(); t++) {
            const char* name = nullptr;
            int p = 0;
            int64_t size = 0;
            check(mc_get_table_info(m_.get(), t, &name, &p, &size));
            (p ? pub : priv).push_back(t);
        }
        if (parameter_type == PUBLIC) return pub;
        if (parameter_type == PRIVATE) return priv;
        pub.insert(pub.end(), priv.begin(), priv.end());
        return pub;
    }
    Evaluation run_eval(const Sessions& s, float* logp, float* full, float* llh) {
        Evaluation e;
        e.n_sessions = s.size();
        if (!s.size()) return e;
        check(mc_evaluate(m_.get(), s.size(), s.pidx.data(), s.clicks.data(),
                          s.length.empty() ? nullptr : s.length.data(), logp, full, llh, &e.llh_sum,
                          e.ppl_log2_sum.data(), e.rank_count.data()));
        return e;
    }
    std::unique_ptr<mc_model, Deleter> m_;
    int64_t n_test_ = 0;
};

}  // namespace mclicks

#endif  // MCLICKS_HPP
