// mc_model.h -- host-side state of one click model on one device, shared by
// the EM (mc_engine.hip) and gradient (mc_grad.hip) translation units.
#ifndef MC_MODEL_H
#define MC_MODEL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"
#include "mc_kernels.h"

#define MC_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                     \
            return MC_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)


struct mc_model {
    int type = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // options
    int reduce_mode = MC_REDUCE_SEGMENTED;
    int item_sessions = 8192;
    int item_window = 1280;
    int sort_by_lcr = 1;
    int block = 256;
    int direct = 0;
    int persistent = 1;
    int packed = 1;           // MC_OPT_PACKED
    int skip_constant = 1;    // MC_OPT_SKIP_CONSTANT
    int fused_merge = 1;      // MC_OPT_FUSED_MERGE
    bool fuse = false;        // the persistent kernel runs the M-step of this layout
    int persist_grid[2] = {0, 0};   // resident workgroups of the persistent window kernel: [0] rank-major, [1] packed items
    // shape
    int T = 1, K = 10, KACC = 10;
    int64_t N = 0, n_qd = 0;
    bool varlen = false;
    bool initialized = false;
    // device buffers
    uint32_t* d_packed = nullptr;   // [n_pk][4] 16-byte records of packed window items' sessions
    uint16_t* d_pidx16 = nullptr;   // [10][n_w - n_pk] window-relative indices of the other window items' sessions
    int32_t* d_pidx = nullptr;      // [10][N - n_w] absolute indices of the other sessions
    int64_t n_pk = 0;               // sessions in packed window items (device positions [0, n_pk))
    int64_t n_w = 0;                // sessions in window items (device positions [0, n_w))
    int64_t n_em_sessions = 0;      // sessions the EM kernels visit (Item::e1)
    uint16_t* d_mask = nullptr;
    uint8_t* d_len = nullptr;
    mc::Item* d_items = nullptr;
    int n_items = 0;
    int n_win_items = 0;      // items [0, n_win_items) have an LDS window (the first
    int n_pk_items = 0;       //   n_pk_items of them packed),
    int n_glob_items = 0;     // then items merged with global atomics,
    int n_dir_items = 0;      // then direct items (single-session groups)
    int max_width = 0;
    int64_t max_item_sessions = 0;
    int32_t* d_shared_idx = nullptr;
    int4* d_split_groups = nullptr;  // per split group: first entry, end entry, slices
    unsigned* d_split_done = nullptr;
    unsigned* d_grid_done = nullptr;
    int64_t n_shared = 0;
    float* d_val[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [buf][t]
    float* d_num[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int2* d_counts = nullptr;
    uint16_t* d_counts16 = nullptr;  // packed counts read by the window flushes
    std::vector<int2> h_counts;   // host copy (occurrences, clicked) per entry
    unsigned long long* d_gacc[2] = {nullptr, nullptr};
    float* d_den_user[2] = {nullptr, nullptr};
    int den_state[3] = {0, 0, 0};  // per table: 0 prior (2), 1 user-set, 2 M-step constant, 3 value only
    float* d_sh_val[2] = {nullptr, nullptr};
    float* d_sh_num[2] = {nullptr, nullptr};
    float* d_sh_den[2] = {nullptr, nullptr};
    double* d_sh_base = nullptr;
    int32_t* d_sh_acc_map = nullptr;
    unsigned long long* d_sh_acc = nullptr;
    int cur = 0;
    bool estep_pending = false;
    size_t lds_bytes = 0;
    size_t mem_bytes = 0;
    std::vector<void*> allocations;
    // ---- gradient path (mc_grad.hip) -------------------------------------
    int n_batches = 1;             // MC_OPT_BATCHES: minibatches of whole query groups
    std::vector<int64_t> batch_sessions;       // sessions per batch
    // Item index lists: for batch b (b = n_batches: every item), window
    // items at win_off[b] (win_cnt[b] of them) and global items at
    // glob_off[b]; split entries of batch b at shared_off[b] in d_shared_idx.
    int32_t* d_lists = nullptr;
    std::vector<int64_t> win_off, win_cnt, glob_off, glob_cnt, shared_off, shared_cnt;
    bool grad_ready = false;
    float* d_logit[2] = {nullptr, nullptr};    // per-qd logits
    float* d_am[2] = {nullptr, nullptr};       // Adam first moments
    float* d_av[2] = {nullptr, nullptr};       // Adam second moments
    float* d_grad[2] = {nullptr, nullptr};     // gradient output of mc_loglik_gradient
    float* d_sh_logit = nullptr;
    float* d_sh_am = nullptr;
    float* d_sh_av = nullptr;
    float* d_sh_grad = nullptr;
    float* d_sh2 = nullptr;                    // alternate Adam state of the shared tables (fused epochs)
    unsigned long long* d_gsh = nullptr;       // shared-table sums of the gradient path, three steps
    double* d_sh_prior = nullptr;              // per shared entry: pseudo-count weight of the prior
    double* d_ll_part = nullptr;               // per item: log-likelihood of its sessions
    int32_t* d_steps = nullptr;                // completed Adam epochs (every batch once)
    int n_cu = 0;
    float adam_lr = 0.05f, adam_b1 = 0.9f, adam_b2 = 0.999f, adam_eps = 1e-8f;
    int adam_prior = 1;
    int adam_next_batch = 0;
    hipGraphExec_t adam_graph = nullptr;       // one epoch (every batch once) of Adam steps
    int adam_graph_first = -1;
    bool value_only = false;                   // tables hold values only (numerators stale)
    // ---- resident held-out set (mc_eval.hip) --------------------------------
    uint32_t* d_test_packed = nullptr;         // [n_test][4] packed test sessions, or null (rank-major set)
    int32_t* d_test_pidx = nullptr;            // [10][n_test] pair indices (-1 past the length / unseen)
    uint16_t* d_test_mask = nullptr;           // [n_test] click bits
    uint8_t* d_test_len = nullptr;             // [n_test] or null (10 ranks)
    int32_t* d_test_perm = nullptr;            // [n_test] caller's index of each stored session
    double* d_test_part = nullptr;             // evaluation partials and totals
    int64_t n_test = 0;
    size_t test_bytes = 0;
    // graph cache, one slot per starting buffer parity
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_iters[2] = {0, 0};
};

namespace mc {

template <class P>
inline mc_status dalloc(mc_model* m, P** ptr, size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? MC_ERR_OOM : MC_ERR_HIP;
    }
    m->allocations.push_back(p);
    m->mem_bytes += bytes;
    *ptr = (P*)p;
    return MC_OK;
}

// Frees every device allocation and cached graph of m.
void release(mc_model* m);
// Fills n floats with v on m's stream (a kernel, safe inside graph capture).
mc_status fill(mc_model* m, float* p, float v, int64_t n);
const char* table_name(int type, int t);
// Resolves, once per device, the device address of the evaluation scratch
// array of mc_evaluate_device (mc_eval.hip); called by mc_model_create with
// the device current.
mc_status eval_scratch_setup(int device);

}  // namespace mc

#endif  // MC_MODEL_H
