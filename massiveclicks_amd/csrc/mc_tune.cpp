// mc_tune.cpp -- autotuner of the EM work decomposition.
//
// The options that shape the E/M kernels (reduction mode, sessions per work
// item and so sessions per lane, LDS window, persistent kernel or one
// workgroup per item, and that workgroup's size) are picked for one model
// type and one training set by measuring graph-replayed EM iterations of
// each candidate on the device: first the reduction mode, then a coordinate
// search over item size, window and launch shape, each step keeping the
// fastest.  The starting point is a candidate, so the result is never slower
// than it (up to measurement noise).

#include <cstdint>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int kNumOptions = MC_N_OPTIONS;

struct Trial {
    int model_type, device;
    int64_t n, n_qd;
    const int32_t *query, *pidx;
    const uint8_t *clicks, *len;
    int iterations;
};

// ms per EM iteration of one option set, or a negative value when the
// options are rejected (e.g. a window too small for the model's layout).
mc_status measure(const Trial& tr, const int64_t* opt, double* ms_per_iter) {
    mc_model* m = nullptr;
    mc_status st = mc_model_create(tr.model_type, tr.device, &m);
    if (st != MC_OK) return st;
    for (int o = 0; o < kNumOptions && st == MC_OK; o++) st = mc_model_set_option(m, o, opt[o]);
    if (st == MC_OK) st = mc_init_parameters(m, tr.n, tr.query, tr.pidx, tr.clicks, tr.len, tr.n_qd);
    double ms = 0.0;
    if (st == MC_OK) st = mc_em_graph_timed(m, tr.iterations, &ms);
    mc_model_destroy(m);
    if (st == MC_ERR_INVALID_ARG || st == MC_ERR_UNSUPPORTED) {
        *ms_per_iter = -1.0;
        return MC_OK;
    }
    if (st != MC_OK) return st;
    *ms_per_iter = ms / tr.iterations;
    return MC_OK;
}

}  // namespace
}  // namespace mc

using namespace mc;

extern "C" mc_status mc_autotune(int model_type, int device, int64_t n, const int32_t* query, const int32_t* pidx,
                                 const uint8_t* clicks, const uint8_t* len, int64_t n_qd, int iterations,
                                 int64_t* options, double* best_ms_per_iter, int* n_trials) {
    if (!query || !pidx || !clicks || !options || iterations < 1) {
        set_error("invalid argument");
        return MC_ERR_INVALID_ARG;
    }
    const Trial tr{model_type, device, n, n_qd, query, pidx, clicks, len, iterations};
    std::vector<int64_t> best(options, options + kNumOptions);
    double best_ms = 0.0;
    int trials = 0;
    mc_status st = measure(tr, best.data(), &best_ms);
    trials++;
    if (st != MC_OK) return st;
    if (best_ms < 0) {
        set_error("the starting options are not valid for this training set");
        return MC_ERR_INVALID_ARG;
    }
    // Tries `value` for option `o`; keeps it when faster.
    auto attempt = [&](int o, int64_t value) -> mc_status {
        if (best[(size_t)o] == value) return MC_OK;
        std::vector<int64_t> cand = best;
        cand[(size_t)o] = value;
        double ms = 0.0;
        mc_status s = measure(tr, cand.data(), &ms);
        trials++;
        if (s != MC_OK) return s;
        if (ms >= 0 && ms < best_ms) {
            best = cand;
            best_ms = ms;
        }
        return MC_OK;
    };
    if ((st = attempt(MC_OPT_REDUCE_MODE, best[MC_OPT_REDUCE_MODE] == MC_REDUCE_SEGMENTED ? MC_REDUCE_ATOMIC
                                                                                            : MC_REDUCE_SEGMENTED)) != MC_OK)
        return st;
    for (int64_t s : {2048, 4096, 8192, 16384})
        if ((st = attempt(MC_OPT_ITEM_SESSIONS, s)) != MC_OK) return st;
    if (best[MC_OPT_REDUCE_MODE] == MC_REDUCE_SEGMENTED) {
        for (int64_t w : {512, 1024, 1280})
            if ((st = attempt(MC_OPT_ITEM_WINDOW, w)) != MC_OK) return st;
        // Launch shape: the persistent kernel (256 lanes) or one workgroup per
        // item of 256 or 512 lanes.
        if ((st = attempt(MC_OPT_PERSISTENT, best[MC_OPT_PERSISTENT] ? 0 : 1)) != MC_OK) return st;
        if (!best[MC_OPT_PERSISTENT])
            for (int64_t b : {256, 512})
                if ((st = attempt(MC_OPT_BLOCK, b)) != MC_OK) return st;
        // Session format of the window items: packed records or rank-major.
        if ((st = attempt(MC_OPT_PACKED, best[MC_OPT_PACKED] ? 0 : 1)) != MC_OK) return st;
    } else {
        for (int64_t b : {256, 512})
            if ((st = attempt(MC_OPT_BLOCK, b)) != MC_OK) return st;
    }
    for (int o = 0; o < kNumOptions; o++) options[o] = best[(size_t)o];
    if (best_ms_per_iter) *best_ms_per_iter = best_ms;
    if (n_trials) *n_trials = trials;
    return MC_OK;
}
