// mc_engine.hip -- kernels and C ABI of the MI355X EM engine.
//
// Per EM iteration two kernels run on the model's stream (the held-out
// evaluation is in mc_eval.hip, the gradient path in mc_grad.hip):
//   em_item_kernel   E-step of one work item per workgroup, per-item
//                    reduction of the expected counts, in-place M-step of the
//                    entries the item owns (parallel_em/kernel.cu:42-89 and
//                    click_models/common.cu:114-192 fused);
//   em_merge_kernel  M-step of entries merged across items and of the shared
//                    (rank-indexed) tables.
// theta_t is read from buffer `cur`, theta_{t+1} written to `cur ^ 1`, so the
// (1,2) reset of the reference (common.cu:71-82, a host fill plus H2D per
// iteration) disappears: the M-step writes prior + sum directly.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"
#include "mc_kernels.h"
#include "mc_layout.h"
#include "mc_model.h"

namespace mc {

namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <int MODEL, int LC, class Src, class Emit>
__device__ __forceinline__ void run_session_len(const int (&pv)[kL], uint32_t m, const Src& src, Emit& em,
                                                const float* sh, float (&acc)[55]) {
    constexpr bool VARLEN = LC < kL;
    const uint32_t mm = m & ((1u << LC) - 1u);
    if (MODEL == kPBM) {
        pbm_session(pv, LC, mm, src, em, sh, acc);
    } else if (MODEL == kUBM) {
        ubm_session(pv, LC, mm, src, em, sh, acc);
    } else if (MODEL == kDBN) {
        if (mm == 0) return;   // every count of a click-free session is a constant
        // The gamma and tau sums go straight into the lane's accumulators.
        dbn_session<VARLEN>(pv, LC, mm, src, em, sh[0], acc[0], acc[1]);
    } else {
        if (mm == 0) return;
        ccm_session<VARLEN>(pv, LC, mm, src, em, sh[0], sh[1], sh[2], acc);
    }
}

// Uneven lists run the session code specialised for their length, so that
// every rank loop and tail index is static; sessions of a work item are
// ordered by length, so a wavefront mostly takes one case.
template <int MODEL, bool VARLEN, class Src, class Emit>
__device__ __forceinline__ void run_session(const int (&pv)[kL], int L, uint32_t m, const Src& src, Emit& em,
                                            const float* sh, float (&acc)[55]) {
    if (!VARLEN) {
        run_session_len<MODEL, kL>(pv, m, src, em, sh, acc);
        return;
    }
    switch (L) {
        case 1: run_session_len<MODEL, 1>(pv, m, src, em, sh, acc); break;
        case 2: run_session_len<MODEL, 2>(pv, m, src, em, sh, acc); break;
        case 3: run_session_len<MODEL, 3>(pv, m, src, em, sh, acc); break;
        case 4: run_session_len<MODEL, 4>(pv, m, src, em, sh, acc); break;
        case 5: run_session_len<MODEL, 5>(pv, m, src, em, sh, acc); break;
        case 6: run_session_len<MODEL, 6>(pv, m, src, em, sh, acc); break;
        case 7: run_session_len<MODEL, 7>(pv, m, src, em, sh, acc); break;
        case 8: run_session_len<MODEL, 8>(pv, m, src, em, sh, acc); break;
        case 9: run_session_len<MODEL, 9>(pv, m, src, em, sh, acc); break;
        default: run_session_len<MODEL, 10>(pv, m, src, em, sh, acc); break;
    }
}

// One lane walks the sessions s0 + tid, s0 + tid + blockDim, ...; the
// indices of the next session are in flight while one is processed.  The
// first session is passed in already loaded.
template <int MODEL, bool VARLEN, bool WIN, class Src, class Emit>
__device__ __forceinline__ void run_item(const Stream& st, const Item& it, const Src& src, Emit& em, const float* sh,
                                         float (&acc)[55], uint32_t m, int L, int (&pv)[kL]) {
    const int nthr = blockDim.x;
    if (VARLEN) {
        // Two-stage prefetch: the length of session s + 2 nthr arrives one
        // session ahead of its indices, which are read only up to it.
        uint32_t mn;
        int Ln;
        load_head(st, it.s0 + (int)threadIdx.x + nthr, it.e1, mn, Ln);
        for (int s = it.s0 + (int)threadIdx.x; s < it.e1; s += nthr) {
            uint32_t m3;
            int L3;
            int pv2[kL];
            load_head(st, s + 2 * nthr, it.e1, m3, L3);
            load_pidx_upto<WIN>(st, s + nthr, it.e1, Ln, pv2);
            run_session<MODEL, VARLEN>(pv, L, m, src, em, sh, acc);
            m = mn;
            L = Ln;
            mn = m3;
            Ln = L3;
#pragma unroll
            for (int r = 0; r < kL; r++) pv[r] = pv2[r];
        }
        return;
    }
    for (int s = it.s0 + (int)threadIdx.x; s < it.e1; s += nthr) {
        uint32_t m2;
        int L2;
        int pv2[kL];
        load_session<VARLEN, WIN>(st, s + nthr, it.e1, m2, L2, pv2);
        run_session<MODEL, VARLEN>(pv, L, m, src, em, sh, acc);
        m = m2;
        L = L2;
#pragma unroll
        for (int r = 0; r < kL; r++) pv[r] = pv2[r];
    }
}

// run_item for packed items: one 16-byte record per session; the next
// session's record is in flight while one is processed (4 registers
// instead of the 12 of a rank-major session).
template <int MODEL, bool VARLEN, class Src, class Emit>
__device__ __forceinline__ void run_item_packed(const Stream& st, const Item& it, const Src& src, Emit& em,
                                                const float* sh, float (&acc)[55], PackedSession cur) {
    const int nthr = blockDim.x;
    for (int s = it.s0 + (int)threadIdx.x; s < it.e1; s += nthr) {
        const PackedSession nxt = load_packed(st, s + nthr, it.e1);
        uint32_t m;
        int L;
        int pv[kL];
        decode_packed<VARLEN>(cur, m, L, pv);
        run_session<MODEL, VARLEN>(pv, L, m, src, em, sh, acc);
        cur = nxt;
    }
}

// Direct items: every session is the only one of its query, so its entries
// receive no other counts; the lane writes theta_{t+1} of its entries itself.
template <int MODEL, bool VARLEN, class Src>
__device__ __forceinline__ void run_item_direct(const KArgs& A, const Stream& st, const Item& it, const Src& src,
                                                const float* sh, float (&acc)[55], uint32_t m, int L,
                                                int (&pv)[kL]) {
    constexpr int T = Shape<MODEL>::T;
    const int nthr = blockDim.x;
    for (int s = it.s0 + (int)threadIdx.x; s < it.e1; s += nthr) {
        uint32_t m2;
        int L2;
        int pv2[kL];
        load_session<VARLEN, false>(st, s + nthr, it.e1, m2, L2, pv2);
        EmitRegs em;
        em.clear();
        run_session<MODEL, VARLEN>(pv, L, m, src, em, sh, acc);
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (!VARLEN || r < L) {
                const int p = pv[r];
                const int2 oc = A.counts[p];
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const float2 bd = base_den<MODEL>(t, oc);
                    const float num = bd.x + em.x[t][r];
                    A.tab[t].val_out[p] = clamp_value(num, bd.y);
                    if (A.write_num) A.tab[t].num_out[p] = num;
                }
            }
        }
        m = m2;
        L = L2;
#pragma unroll
        for (int r = 0; r < kL; r++) pv[r] = pv2[r];
    }
}

// Cross-workgroup hand-over inside one launch (fused M-step).  The data
// handed over (fixed-point sums, counters) is only ever written with
// device-scope atomics, which are performed at the device's coherence point;
// so the writer only has to wait until its atomics have completed before it
// bumps a counter (an agent-scope release fence would also write back the
// XCD's L2 for plain stores, which nothing here needs and which cost 4x the
// kernel's time), and the reader reads with device-coherent loads.
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }

__device__ __forceinline__ unsigned long long load_acc(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// M-step of the shared (rank-indexed) tables from their fixed-point sums and
// FP64 bases, threads tid < K; then the accumulators are cleared for the next
// iteration.  COHERENT: the sums were added during this launch.
template <bool COHERENT>
__device__ __forceinline__ void finalize_shared(const KArgs& A, int tid, int nthr) {
    const int k = tid;
    if (k < A.K) {
        const int an = A.sh_acc_map[2 * k], ad = A.sh_acc_map[2 * k + 1];
        const long long tn = an >= 0 ? (long long)(COHERENT ? load_acc(&A.sh_acc[an]) : A.sh_acc[an]) : 0;
        const long long td = ad >= 0 ? (long long)(COHERENT ? load_acc(&A.sh_acc[ad]) : A.sh_acc[ad]) : 0;
        const float num = (float)(A.sh_base[2 * k] + (double)tn * kFixInv);
        const float den = (float)(A.sh_base[2 * k + 1] + (double)td * kFixInv);
        A.sh_num_out[k] = num;
        A.sh_den_out[k] = den;
        A.sh_val_out[k] = clamp_value(num, den);
    }
    __syncthreads();
    for (int j = tid; j < A.kacc; j += nthr) A.sh_acc[j] = 0ull;
}

// M-step of a per-qd entry whose sums several work items added to gacc; the
// accumulators are cleared for the next iteration.
template <int MODEL, bool COHERENT>
__device__ __forceinline__ void finalize_merged_entry(const KArgs& A, int p) {
    constexpr int T = Shape<MODEL>::T;
    const int2 oc = A.counts[p];
#pragma unroll
    for (int t = 0; t < T; t++) {
        const float2 bd = base_den<MODEL>(t, oc);
        const unsigned long long v = COHERENT ? load_acc(&A.tab[t].gacc[p]) : A.tab[t].gacc[p];
        const float num = fix_to_float(bd.x, v);
        A.tab[t].val_out[p] = clamp_value(num, bd.y);
        if (A.write_num) A.tab[t].num_out[p] = num;
        A.tab[t].gacc[p] = 0ull;
    }
}

// Fused M-step of a split query group after a workgroup has added one of
// its slices: the workgroup that completes the group finalises its entries.
// Block-uniform call (it.split >= 0); `flag` is a shared int.
template <int MODEL>
__device__ __forceinline__ void fused_split_merge(const KArgs& A, const Item& it, int tid, int nthr, int& flag) {
    wait_memory();
    __syncthreads();
    const int4 g = A.split_groups[it.split];
    if (tid == 0) {
        flag = atomicAdd(&A.split_done[it.split], 1u) == (unsigned)g.z - 1u;
        wait_memory();
    }
    __syncthreads();
    if (flag) {
        wait_memory();
        for (int p = g.x + tid; p < g.y; p += nthr) finalize_merged_entry<MODEL, true>(A, p);
        if (tid == 0) A.split_done[it.split] = 0u;
    }
}

// Fused M-step of the shared tables: after a workgroup's shared-table sums
// are in, the last of the launch's n_groups workgroups finalises the tables.
__device__ __forceinline__ void fused_shared_merge(const KArgs& A, int tid, int nthr, unsigned n_groups, int& flag) {
    wait_memory();
    __syncthreads();
    if (tid == 0) {
        flag = atomicAdd(A.grid_done, 1u) == n_groups - 1u;
        wait_memory();
    }
    __syncthreads();
    if (flag) {
        wait_memory();
        finalize_shared<true>(A, tid, nthr);
        if (tid == 0) *A.grid_done = 0u;
    }
}

// Session paths, one kernel instantiation each (each gets its own register
// allocation): window items in LDS (rank-major or packed sessions), items
// without a window (global 64-bit fixed-point atomics), and the float-atomic
// mode.
enum { kPathLds = 0, kPathGlobal = 1, kPathAtomic = 2, kPathDirect = 3, kPathLdsFloat = 4, kPathPacked = 5 };

// E-step and in-item M-step of one work item (item index item_offset +
// blockIdx.x).  Dynamic LDS (window path): width x T fixed-point
// accumulators, then width x T staged values of theta_t (both interleaved per
// entry).  kPathLdsFloat (float-atomic mode) pre-reduces in float LDS
// accumulators and adds every window entry to global float accumulators,
// which the merge kernel finalises.
template <int MODEL, bool VARLEN, int PATH>
__global__ void __launch_bounds__(512) em_item_kernel(KArgs A, int item_offset) {
    constexpr int T = Shape<MODEL>::T;
    constexpr int K = Shape<MODEL>::K;
    constexpr int KACC = Shape<MODEL>::KACC;
    extern __shared__ unsigned long long lds[];
    __shared__ float sh[K < 64 ? 64 : K];
    __shared__ long long red[KACC][8];
    __shared__ int last_one;

    const Item it = A.items[item_offset + blockIdx.x];
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int W = it.width;
    constexpr bool fwin = PATH == kPathLdsFloat;
    constexpr bool packed = PATH == kPathPacked;
    constexpr bool fixwin = PATH == kPathLds || packed;   // fixed-point LDS window
    constexpr bool win = fixwin || fwin;
    unsigned long long* lacc = lds;
    float* facc = (float*)lds;
    // Fixed-point windows take max_width slots per table (copies of the
    // entries of narrow windows), then the staged values.
    const int Wmax = A.max_width;
    const AccWindow<T> aw{lacc, Wmax, fixwin ? acc_copies_log2(W, Wmax) : 0};
    const int WS = W << aw.rsh;
    float* lval = fwin ? facc + T * W : (float*)(lds + T * Wmax);
    // UBM examination slots in LDS after the window (fixed-point window path).
    constexpr bool lslots = MODEL == kUBM && fixwin;
    unsigned long long* slots = lds + ((size_t)T * Wmax * 12 + 7) / 8;

    // First session's loads go out before the window is staged.
    const Stream st = make_stream(A);
    uint32_t m0 = 0;
    int L0 = 0;
    int pv0[kL];
    PackedSession r0;
    if (packed) r0 = load_packed(st, it.s0 + tid, it.e1);
    else load_session<VARLEN, win>(st, it.s0 + tid, it.e1, m0, L0, pv0);
    for (int k = tid; k < K; k += nthr) sh[k] = A.sh_val_in[k];
    if (win) {
        for (int i = tid; i < W; i += nthr) {
#pragma unroll
            for (int t = 0; t < T; t++) {
                if (fwin) facc[i * T + t] = 0.f;
                lval[i * T + t] = A.tab[t].val_in[it.base + i];
            }
        }
        if (!fwin)
            for (int i = tid; i < WS; i += nthr)
#pragma unroll
                for (int t = 0; t < T; t++) lacc[aw.slot(t, i)] = 0ull;
    }
    if (lslots)
        for (int i = tid; i < kUbmSlots * kUbmCopies; i += nthr) slots[i] = 0ull;
    __syncthreads();

    float acc[55];
#pragma unroll
    for (int k = 0; k < 55; k++) acc[k] = 0.f;

    const SrcGlobal gsrc{A.tab[0].val_in, A.tab[1].val_in};
    if (PATH == kPathAtomic) {
        EmitGlobalFloat em{A.tab[0].gaccf, A.tab[1].gaccf};
        run_item<MODEL, VARLEN, false>(st, it, gsrc, em, sh, acc, m0, L0, pv0);
    } else if (fixwin) {
        EmitLds<T> em{aw, tid & ((1 << aw.rsh) - 1), slots + (tid & (kUbmCopies - 1))};
        const SrcLds<T> lsrc{lval};
        if (packed) run_item_packed<MODEL, VARLEN>(st, it, lsrc, em, sh, acc, r0);
        else run_item<MODEL, VARLEN, true>(st, it, lsrc, em, sh, acc, m0, L0, pv0);
    } else if (PATH == kPathLdsFloat) {
        EmitLdsFloat<T> em{facc};
        const SrcLds<T> lsrc{lval};
        run_item<MODEL, VARLEN, true>(st, it, lsrc, em, sh, acc, m0, L0, pv0);
    } else if (PATH == kPathDirect) {
        run_item_direct<MODEL, VARLEN>(A, st, it, gsrc, sh, acc, m0, L0, pv0);
    } else {
        EmitGlobalFix em{A.tab[0].gacc, A.tab[1].gacc};
        run_item<MODEL, VARLEN, false>(st, it, gsrc, em, sh, acc, m0, L0, pv0);
    }

    // Shared tables: fixed point, wave sums, across waves, one atomic per
    // accumulator and work item (integer adds: order-independent).
    const int wave = tid >> 6, lane = tid & 63, nwave = nthr >> 6;
#pragma unroll
    for (int k = 0; k < KACC; k++) {
        long long v = wave_sum(to_fix(acc[k]));
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    for (int k = tid; k < KACC; k += nthr) {
        long long v = 0;
        for (int w = 0; w < nwave; w++) v += red[k][w];
        if (lslots)
            for (int c = 0; c < kUbmCopies; c++) v += (long long)slots[k * kUbmCopies + c];
        if (v != 0) atomicAdd(&A.sh_acc[k], (unsigned long long)v);
    }

    if (fwin) {
        for (int i = tid; i < W; i += nthr) {
#pragma unroll
            for (int t = 0; t < T; t++) {
                const float v = facc[i * T + t];
                if (v != 0.f) unsafeAtomicAdd(&A.tab[t].gaccf[it.base + i], v);
            }
        }
    }
    // Flush the LDS window, four entries per thread per pass with the
    // constant loads issued together.
    if (win && !fwin) {
        const bool interior = (it.flags & kItemInterior) != 0;
        for (int i0 = 0; i0 < W; i0 += 4 * nthr) {
            uint32_t pk[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * nthr + tid;
                pk[u] = (interior && i < W) ? A.counts16[it.base + i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * nthr + tid;
                if (i >= W) continue;
                const int p = it.base + i;
                const int2 oc = interior ? unpack_counts(pk[u], A.counts, p) : make_int2(0, 0);
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const unsigned long long v = aw.sum(t, i);
                    if (interior) {
                        const float2 bd = base_den<MODEL>(t, oc);
                        const float num = fix_to_float(bd.x, v);
                        A.tab[t].val_out[p] = clamp_value(num, bd.y);
                        if (A.write_num) A.tab[t].num_out[p] = num;
                    } else if (v != 0ull) {
                        atomicAdd(&A.tab[t].gacc[p], v);
                    }
                }
            }
        }
    }
    // Fused M-step (the launch is the iteration's only E-step launch): split
    // groups, then the shared tables, as in the persistent kernel.
    if (A.fuse_merge) {
        if (it.split >= 0) fused_split_merge<MODEL>(A, it, tid, nthr, last_one);
        fused_shared_merge(A, tid, nthr, gridDim.x, last_one);
    }
}

// Persistent form of the window path over items [first, first + n_win), all
// packed (PACKED) or all rank-major: workgroup b of G resident ones
// processes one window item per round, item kG + b in even rounds k and
// kG + G - 1 - b in odd ones.  Items are ordered largest first, so this
// back-and-forth order pairs the larger items of one round with the smaller
// of the next, while the items of a round stay adjacent in memory; a fixed
// assignment keeps the per-lane float sums, and so the results, identical
// from run to run.
// While item k is processed, the window values of the workgroup's next item
// and its first session are in flight (registers), and the constants of
// item k's entries arrive during its session loop, so the round trips of
// staging and flushing overlap with the E-step and the flush.  The shared-table counts stay
// in registers across items and are reduced once per workgroup.  Requires
// blockDim.x == 256 and width <= 256 * PER.
// Waves per SIMD the persistent kernel asks the compiler to fit: ten-rank DBN
// goes from 4 waves (108 VGPRs) to 5 (89-95 VGPRs, no spills), 1-3 %
// faster (at 6, 80 VGPRs with spills, it is 8 % slower and its 1280-entry
// window's LDS allows only 5 workgroups per CU anyway); ten-rank CCM over
// packed sessions runs 6 (80 VGPRs, 5 dwords spilled; one per-qd table, so
// 6 windows fit), 2-3 % faster than 5 (its rank-major form would spill 15
// dwords and keeps 5); the uneven-length UBM kernel goes from 2 waves (181 VGPRs)
// to 3 (168 VGPRs, 4 dwords spilled), 22 % faster.  Ten-rank PBM fits 7
// without a hint (the host caps it at 6 per CU: 7 measured slower), and
// ten-rank UBM fits 6 (79 VGPRs) but its window plus examination slots
// allow 5 workgroups per CU.
template <int MODEL, bool VARLEN, bool PACKED>
struct PersistWaves {
    static constexpr int value = (!VARLEN && MODEL == kDBN)   ? 5
                                 : (!VARLEN && MODEL == kCCM) ? (PACKED ? 6 : 5)
                                 : (VARLEN && MODEL == kUBM)  ? 3
                                                              : 1;
};

template <int MODEL, bool VARLEN, int PER, bool PACKED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PersistWaves<MODEL, VARLEN, PACKED>::value)))
em_window_persistent_kernel(KArgs A, int first, int n_win) {
    constexpr int T = Shape<MODEL>::T;
    constexpr int K = Shape<MODEL>::K;
    constexpr int KACC = Shape<MODEL>::KACC;
    constexpr int NT = 256;
    extern __shared__ unsigned long long lds[];
    __shared__ float sh[K < 64 ? 64 : K];
    __shared__ long long red[KACC][4];
    __shared__ int last_one;

    const int tid = threadIdx.x;
    const int Wmax = A.max_width;
    unsigned long long* lacc = lds;
    float* lval = (float*)(lds + T * Wmax);
    // UBM examination slots in LDS after the window, kept across items.
    constexpr bool lslots = MODEL == kUBM;
    unsigned long long* slots = lds + ((size_t)T * Wmax * 12 + 7) / 8;
    if (lslots)
        for (int i = tid; i < kUbmSlots * kUbmCopies; i += NT) slots[i] = 0ull;
    const Stream st = make_stream(A);
    for (int k = tid; k < K; k += NT) sh[k] = A.sh_val_in[k];

    const int G = (int)gridDim.x;
    auto item_of_round = [&](int k) { return k * G + ((k & 1) ? G - 1 - (int)blockIdx.x : (int)blockIdx.x); };
    int round = 0;
    int cur = item_of_round(0);
    Item it = cur < n_win ? A.items[first + cur] : Item{};

    // Window values of the current item, and its first session.
    float wv[PER][T];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = u * NT + tid;
#pragma unroll
        for (int t = 0; t < T; t++) wv[u][t] = (i < it.width) ? A.tab[t].val_in[it.base + i] : 0.f;
    }
    uint32_t m0 = 0;
    int L0 = 0;
    int pv0[kL];
    PackedSession r0;
    if (PACKED) r0 = load_packed(st, it.s0 + tid, it.e1);
    else load_session<VARLEN, true>(st, it.s0 + tid, it.e1, m0, L0, pv0);

    float acc[55];
#pragma unroll
    for (int k = 0; k < 55; k++) acc[k] = 0.f;

    while (cur < n_win) {
        const int W = it.width;
        const AccWindow<T> aw{lacc, Wmax, acc_copies_log2(W, Wmax)};
        const int WS = W << aw.rsh;
        // Stage the window, take the next item.
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = u * NT + tid;
            if (i < W) {
#pragma unroll
                for (int t = 0; t < T; t++) lval[i * T + t] = wv[u][t];
            }
            if (i < WS) {
#pragma unroll
                for (int t = 0; t < T; t++) lacc[aw.slot(t, i)] = 0ull;
            }
        }
        __syncthreads();
        const int nxt = item_of_round(round + 1);
        const Item itn = nxt < n_win ? A.items[first + nxt] : Item{};
        // In flight during the session loop: next window values, this
        // window's counts.
        // Ten-rank kernels load the next window after the session loop
        // instead (its latency then overlaps the flush), which frees the
        // window registers during the loop: fewer spills at 5 waves per SIMD,
        // 1-3 % faster; the uneven-length kernels measured 1-3 % slower so.
        constexpr bool LATE_WINDOW = !VARLEN;
        uint32_t cnt[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = u * NT + tid;
            if (!LATE_WINDOW) {
#pragma unroll
                for (int t = 0; t < T; t++) wv[u][t] = (i < itn.width) ? A.tab[t].val_in[itn.base + i] : 0.f;
            }
            cnt[u] = (i < W) ? A.counts16[it.base + i] : 0u;
        }
        {
            EmitLds<T> em{aw, tid & ((1 << aw.rsh) - 1), slots + (tid & (kUbmCopies - 1))};
            const SrcLds<T> lsrc{lval};
            if (PACKED) run_item_packed<MODEL, VARLEN>(st, it, lsrc, em, sh, acc, r0);
            else run_item<MODEL, VARLEN, true>(st, it, lsrc, em, sh, acc, m0, L0, pv0);
        }
        const bool interior = (it.flags & kItemInterior) != 0;
        // Full counts of the entries with 255 or more occurrences, loaded
        // before the prefetches below and without a branch (lanes without
        // such an entry read the item's first entry, a line the wave
        // shares): a conditional load per entry made the flush wait, entry
        // by entry, for everything in flight, the prefetches and the
        // previous entry's stores included (PBM and UBM 5 % slower).
        int2 full[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = u * NT + tid;
            full[u] = A.counts[cnt[u] == kCountsEscape ? it.base + i : it.base];
        }
        if (LATE_WINDOW) {
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int i = u * NT + tid;
#pragma unroll
                for (int t = 0; t < T; t++) wv[u][t] = (i < itn.width) ? A.tab[t].val_in[itn.base + i] : 0.f;
            }
        }
        // First session of the next item goes out before the flush.
        if (PACKED) r0 = load_packed(st, itn.s0 + tid, itn.e1);
        else load_session<VARLEN, true>(st, itn.s0 + tid, itn.e1, m0, L0, pv0);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = u * NT + tid;
            if (i >= W) continue;
            const int p = it.base + i;
            const int2 ocu = interior ? unpack_counts(cnt[u], full[u]) : make_int2(0, 0);
#pragma unroll
            for (int t = 0; t < T; t++) {
                const unsigned long long v = aw.sum(t, i);
                if (interior) {
                    const float2 bd = base_den<MODEL>(t, ocu);
                    const float num = fix_to_float(bd.x, v);
                    A.tab[t].val_out[p] = clamp_value(num, bd.y);
                    if (A.write_num) A.tab[t].num_out[p] = num;
                } else if (v != 0ull) {
                    atomicAdd(&A.tab[t].gacc[p], v);
                }
            }
        }
        // Fused M-step of a split query group: the workgroup that adds the
        // group's last slice finalises the group's entries.
        if (A.fuse_merge && it.split >= 0) fused_split_merge<MODEL>(A, it, tid, NT, last_one);
        __syncthreads();
        round++;
        cur = nxt;
        it = itn;
    }

    // Shared tables: one reduction and one atomic per accumulator per workgroup.
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int k = 0; k < KACC; k++) {
        long long v = wave_sum(to_fix(acc[k]));
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    for (int k = tid; k < KACC; k += NT) {
        long long v = red[k][0] + red[k][1] + red[k][2] + red[k][3];
        if (lslots)
            for (int c = 0; c < kUbmCopies; c++) v += (long long)slots[k * kUbmCopies + c];
        if (v != 0) atomicAdd(&A.sh_acc[k], (unsigned long long)v);
    }
    // Fused M-step of the shared tables: the last workgroup to finish.
    if (A.fuse_merge) fused_shared_merge(A, tid, NT, (unsigned)G, last_one);
}

// Merge kernel: block 0 finalises the shared tables, the other blocks
// finalise per-qd entries accumulated in global memory (entries merged across
// items in segmented mode, every entry in atomic mode).
template <int MODEL>
__global__ void __launch_bounds__(256) em_merge_kernel(KArgs A) {
    constexpr int T = Shape<MODEL>::T;
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        finalize_shared<false>(A, tid, (int)blockDim.x);
        return;
    }
    const int64_t i = (int64_t)(blockIdx.x - 1) * blockDim.x + tid;
    if (A.mode == 1) {
        if (i >= A.n_qd) return;
        const int2 oc = A.counts[i];
#pragma unroll
        for (int t = 0; t < T; t++) {
            const float2 bd = base_den<MODEL>(t, oc);
            const float v = A.tab[t].gaccf[i];
            const float num = (float)((double)bd.x + (double)v);
            A.tab[t].val_out[i] = clamp_value(num, bd.y);
            if (A.write_num) A.tab[t].num_out[i] = num;
            A.tab[t].gaccf[i] = 0.f;
        }
    } else {
        if (i >= A.n_shared) return;
        finalize_merged_entry<MODEL, false>(A, A.shared_idx[i]);
    }
}

// Fills a float array with a constant (used instead of memsets so that the
// same code is safe inside captured graphs).
__global__ void fill_kernel(float* p, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace mc

// ===========================================================================
// Host side
// ===========================================================================
using namespace mc;

namespace mc {

const char* table_name(int type, int t) {
    switch (type) {
        case MC_PBM: return t == 0 ? "attractiveness" : "examination";
        case MC_UBM: return t == 0 ? "attractiveness" : "examination";
        case MC_CCM: return t == 0 ? "attractiveness" : "continuation";
        case MC_DBN: return t == 0 ? "attractiveness" : (t == 1 ? "satisfaction" : "continuation");
    }
    return "?";
}

void release(mc_model* m) {
    if (m->adam_graph) { (void)hipGraphExecDestroy(m->adam_graph); m->adam_graph = nullptr; }
    m->grad_ready = false;
    for (int b = 0; b < 2; b++)
        if (m->graph[b]) { (void)hipGraphExecDestroy(m->graph[b]); m->graph[b] = nullptr; }
    for (void* p : m->allocations) (void)hipFree(p);
    m->allocations.clear();
    m->mem_bytes = 0;
    m->d_test_packed = nullptr;
    m->d_test_pidx = nullptr;
    m->d_test_mask = nullptr;
    m->d_test_len = nullptr;
    m->d_test_perm = nullptr;
    m->d_test_part = nullptr;
    m->n_test = 0;
    m->test_bytes = 0;
    m->initialized = false;
}

mc_status fill(mc_model* m, float* p, float v, int64_t n) {
    if (n <= 0) return MC_OK;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, p, v, n);
    MC_HIP(hipGetLastError());
    return MC_OK;
}

}  // namespace mc

namespace {

void shape_of(int type, int* T, int* K, int* KACC) {
    switch (type) {
        case MC_PBM: *T = 1; *K = 10; *KACC = 10; break;
        case MC_UBM: *T = 1; *K = 55; *KACC = 55; break;
        case MC_DBN: *T = 2; *K = 1; *KACC = 2; break;
        default: *T = 1; *K = 3; *KACC = 6; break;
    }
}

KArgs make_args(mc_model* m, int cur, int write_num) {
    KArgs A;
    std::memset(&A, 0, sizeof(A));
    A.packed = m->d_packed;
    A.pidx16 = m->d_pidx16;
    A.pidx = m->d_pidx;
    A.mask = m->d_mask;
    A.len = m->varlen ? m->d_len : nullptr;
    A.N = m->N;
    A.n_w = m->n_w;
    A.n_pk = m->n_pk;
    A.items = m->d_items;
    A.n_items = m->n_items;
    A.mode = m->reduce_mode;
    for (int t = 0; t < 2; t++) {
        int tt = t < m->T ? t : 0;
        A.tab[t].val_in = m->d_val[cur][tt];
        A.tab[t].val_out = m->d_val[cur ^ 1][tt];
        A.tab[t].num_out = m->d_num[cur ^ 1][tt];
        A.tab[t].gacc = m->d_gacc[tt];
        A.tab[t].gaccf = (float*)m->d_gacc[tt];
    }
    A.sh_val_in = m->d_sh_val[cur];
    A.sh_acc = m->d_sh_acc;
    A.shared_idx = m->d_shared_idx;
    A.n_shared = m->n_shared;
    A.n_qd = m->n_qd;
    A.n_tables = m->T;
    A.write_num = write_num;
    A.counts = m->d_counts;
    A.counts16 = m->d_counts16;
    A.max_width = m->max_width;
    A.K = m->K;
    A.kacc = m->KACC;
    A.sh_base = m->d_sh_base;
    A.sh_acc_map = m->d_sh_acc_map;
    A.sh_val_out = m->d_sh_val[cur ^ 1];
    A.sh_num_out = m->d_sh_num[cur ^ 1];
    A.sh_den_out = m->d_sh_den[cur ^ 1];
    A.fuse_merge = m->fuse ? 1 : 0;
    A.split_groups = m->d_split_groups;
    A.split_done = m->d_split_done;
    A.grid_done = m->d_grid_done;
    return A;
}

template <int MODEL, int PATH>
void launch_path(mc_model* m, const KArgs& A, int first, int count, size_t lds) {
    if (count <= 0) return;
    if (m->varlen)
        hipLaunchKernelGGL((em_item_kernel<MODEL, true, PATH>), dim3(count), dim3(m->block), lds, m->stream, A, first);
    else
        hipLaunchKernelGGL((em_item_kernel<MODEL, false, PATH>), dim3(count), dim3(m->block), lds, m->stream, A, first);
}

// Window width covered by the persistent kernel's register staging.
constexpr int kPersistPer = 5;

// Resident workgroups per CU of the model's persistent window kernel.
template <int MODEL, bool PACKED>
hipError_t persistent_occupancy_of(const mc_model* m, int* per_cu) {
    return m->varlen
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     per_cu, em_window_persistent_kernel<MODEL, true, kPersistPer, PACKED>, 256, m->lds_bytes)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     per_cu, em_window_persistent_kernel<MODEL, false, kPersistPer, PACKED>, 256, m->lds_bytes);
}

template <bool PACKED>
hipError_t persistent_occupancy_pk(const mc_model* m, int* per_cu) {
    switch (m->type) {
        case MC_PBM: return persistent_occupancy_of<kPBM, PACKED>(m, per_cu);
        case MC_UBM: return persistent_occupancy_of<kUBM, PACKED>(m, per_cu);
        case MC_DBN: return persistent_occupancy_of<kDBN, PACKED>(m, per_cu);
        default: return persistent_occupancy_of<kCCM, PACKED>(m, per_cu);
    }
}

hipError_t persistent_occupancy(const mc_model* m, bool packed, int* per_cu) {
    return packed ? persistent_occupancy_pk<true>(m, per_cu) : persistent_occupancy_pk<false>(m, per_cu);
}

template <int MODEL, bool PACKED>
void launch_persistent(mc_model* m, const KArgs& A, int first, int count) {
    if (count <= 0) return;
    const dim3 grid((unsigned)std::min(count, m->persist_grid[PACKED]));
    if (m->varlen)
        hipLaunchKernelGGL((em_window_persistent_kernel<MODEL, true, kPersistPer, PACKED>), grid, dim3(256),
                           m->lds_bytes, m->stream, A, first, count);
    else
        hipLaunchKernelGGL((em_window_persistent_kernel<MODEL, false, kPersistPer, PACKED>), grid, dim3(256),
                           m->lds_bytes, m->stream, A, first, count);
}

template <int MODEL>
void launch_item(mc_model* m, const KArgs& A) {
    if (m->reduce_mode == MC_REDUCE_ATOMIC) {
        // LDS pre-reduction of window items, then float atomics; items
        // without a window add straight to the global accumulators.
        launch_path<MODEL, kPathLdsFloat>(m, A, 0, m->n_win_items, m->lds_bytes);
        launch_path<MODEL, kPathAtomic>(m, A, m->n_win_items, m->n_glob_items, 0);
    } else {
        const int n_pk = m->n_pk_items, n_rm = m->n_win_items - m->n_pk_items;
        if (m->persistent) {
            launch_persistent<MODEL, true>(m, A, 0, n_pk);
            launch_persistent<MODEL, false>(m, A, n_pk, n_rm);
        } else {
            launch_path<MODEL, kPathPacked>(m, A, 0, n_pk, m->lds_bytes);
            launch_path<MODEL, kPathLds>(m, A, n_pk, n_rm, m->lds_bytes);
        }
        launch_path<MODEL, kPathGlobal>(m, A, m->n_win_items, m->n_glob_items, 0);
        launch_path<MODEL, kPathDirect>(m, A, m->n_win_items + m->n_glob_items, m->n_dir_items, 0);
    }
}

mc_status launch_estep(mc_model* m, int cur, int write_num) {
    KArgs A = make_args(m, cur, write_num);
    switch (m->type) {
        case MC_PBM: launch_item<kPBM>(m, A); break;
        case MC_UBM: launch_item<kUBM>(m, A); break;
        case MC_DBN: launch_item<kDBN>(m, A); break;
        default: launch_item<kCCM>(m, A); break;
    }
    MC_HIP(hipGetLastError());
    return MC_OK;
}

mc_status launch_mstep(mc_model* m, int cur, int write_num) {
    if (m->fuse) return MC_OK;   // done by the persistent kernel
    KArgs A = make_args(m, cur, write_num);
    int64_t n_entries = m->reduce_mode == MC_REDUCE_ATOMIC ? m->n_qd : m->n_shared;
    unsigned grid = (unsigned)(1 + (n_entries + 255) / 256);
    switch (m->type) {
        case MC_PBM: hipLaunchKernelGGL(em_merge_kernel<kPBM>, dim3(grid), dim3(256), 0, m->stream, A); break;
        case MC_UBM: hipLaunchKernelGGL(em_merge_kernel<kUBM>, dim3(grid), dim3(256), 0, m->stream, A); break;
        case MC_DBN: hipLaunchKernelGGL(em_merge_kernel<kDBN>, dim3(grid), dim3(256), 0, m->stream, A); break;
        default: hipLaunchKernelGGL(em_merge_kernel<kCCM>, dim3(grid), dim3(256), 0, m->stream, A); break;
    }
    MC_HIP(hipGetLastError());
    return MC_OK;
}

// Captures n_iter iterations starting from buffer parity `cur` into a graph
// (cached per parity).
mc_status get_graph(mc_model* m, int n_iter, int cur, hipGraphExec_t* out) {
    if (m->graph[cur] && m->graph_iters[cur] == n_iter) { *out = m->graph[cur]; return MC_OK; }
    if (m->graph[cur]) { MC_HIP(hipGraphExecDestroy(m->graph[cur])); m->graph[cur] = nullptr; }
    hipGraph_t g;
    MC_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    int c = cur;
    mc_status st = MC_OK;
    for (int i = 0; i < n_iter && st == MC_OK; i++) {
        const int wn = i == n_iter - 1;   // numerators only for the final state
        st = launch_estep(m, c, wn);
        if (st == MC_OK) st = launch_mstep(m, c, wn);
        c ^= 1;
    }
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    if (st != MC_OK) return st;
    MC_HIP(e);
    e = hipGraphInstantiate(&m->graph[cur], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MC_HIP(e);
    m->graph_iters[cur] = n_iter;
    *out = m->graph[cur];
    return MC_OK;
}


}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* mc_last_error(void) { return g_last_error.c_str(); }

mc_status mc_model_create(int model_type, int device, mc_model** out) {
    MC_REQUIRE(out, "null output pointer");
    MC_REQUIRE(model_type >= 0 && model_type <= 3, "model_type must be 0 (PBM), 1 (CCM), 2 (DBN) or 3 (UBM)");
    int n_dev = 0;
    MC_HIP(hipGetDeviceCount(&n_dev));
    MC_REQUIRE(device >= 0 && device < n_dev, "no such HIP device");
    MC_HIP(hipSetDevice(device));
    mc_model* m = new mc_model;
    m->type = model_type;
    m->device = device;
    shape_of(model_type, &m->T, &m->K, &m->KACC);
    const mc_status ss = eval_scratch_setup(device);
    if (ss != MC_OK) {
        delete m;
        return ss;
    }
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete m;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return MC_ERR_HIP;
    }
    *out = m;
    return MC_OK;
}

mc_status mc_get_model_type(const mc_model* m, int* model_type) {
    MC_REQUIRE(m && model_type, "null argument");
    *model_type = m->type;
    return MC_OK;
}

mc_status mc_model_set_option(mc_model* m, int option, int64_t value) {
    MC_REQUIRE(m, "null model");
    MC_REQUIRE(!m->initialized, "options must be set before mc_init_parameters");
    switch (option) {
        case MC_OPT_REDUCE_MODE:
            MC_REQUIRE(value == MC_REDUCE_SEGMENTED || value == MC_REDUCE_ATOMIC, "bad reduce mode");
            m->reduce_mode = (int)value;
            break;
        case MC_OPT_ITEM_SESSIONS:
            MC_REQUIRE(value >= 64 && value <= (1 << 20), "item sessions out of range [64, 2^20]");
            m->item_sessions = (int)value;
            break;
        case MC_OPT_ITEM_WINDOW:
            MC_REQUIRE(value >= 0 && value <= 16384, "item window out of range [0, 16384]");
            m->item_window = (int)value;
            break;
        case MC_OPT_SORT_BY_LCR:
            m->sort_by_lcr = value != 0;
            break;
        case MC_OPT_DIRECT:
            m->direct = value != 0;
            break;
        case MC_OPT_PERSISTENT:
            m->persistent = value != 0;
            break;
        case MC_OPT_PACKED:
            m->packed = value != 0;
            break;
        case MC_OPT_SKIP_CONSTANT:
            m->skip_constant = value != 0;
            break;
        case MC_OPT_FUSED_MERGE:
            m->fused_merge = value != 0;
            break;
        case MC_OPT_BATCHES:
            MC_REQUIRE(value >= 1 && value <= 4096, "batches out of range [1, 4096]");
            m->n_batches = (int)value;
            break;
        case MC_OPT_BLOCK:
            MC_REQUIRE(value == 64 || value == 128 || value == 256 || value == 512, "block must be 64/128/256/512");
            m->block = (int)value;
            break;
        default:
            set_error("unknown option");
            return MC_ERR_INVALID_ARG;
    }
    return MC_OK;
}

mc_status mc_init_parameters(mc_model* m, int64_t n, const int32_t* query, const int32_t* pidx,
                             const uint8_t* clicks, const uint8_t* len, int64_t n_qd) {
    MC_REQUIRE(m && query && pidx && clicks, "null argument");
    // The session stream is addressed with 32-bit buffer offsets: 20 B per
    // session of a window item (checked here), 40 B per other session
    // (checked by the planner, which decides which sessions those are).
    MC_REQUIRE(n >= 1 && n * 20 < (int64_t)UINT32_MAX, "session count out of range (at most 214,748,364)");
    MC_REQUIRE(n_qd >= 1 && n_qd <= INT32_MAX - 1, "n_qd out of range");
    MC_HIP(hipSetDevice(m->device));
    if (m->initialized) release(m);
    const int T = m->T, K = m->K, KACC = m->KACC;
    m->N = n;
    m->n_qd = n_qd;
    m->varlen = len != nullptr;

    HostLayout h;
    try {
        MC_TRY(plan_layout(m, n, query, pidx, clicks, len, n_qd, &h));
    } catch (std::bad_alloc&) {
        set_error("out of host memory while planning the device layout");
        return MC_ERR_OOM;
    }
    const int max_width = m->max_width;
    const std::vector<Item>& items = h.items;
    const std::vector<int32_t>& shared_idx = h.shared_idx;
    const std::vector<int32_t>& lists = h.lists;
    const std::vector<int2>& counts = h.counts;
    const std::vector<double>& sh_base = h.sh_base;
    const std::vector<int32_t>& sh_map = h.sh_map;
    const std::vector<uint16_t>& h_pidx16 = h.pidx16;
    const std::vector<int32_t>& h_pidx = h.pidx32;
    const std::vector<uint16_t>& h_mask = h.mask;
    const std::vector<uint8_t>& h_len = h.len;
    {
        // The window (plus the kernels' few KB of static LDS) must fit one
        // workgroup's LDS.
        int max_lds = 0;
        MC_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device));
        MC_REQUIRE(m->lds_bytes + 4096 <= (size_t)max_lds,
                   "item window too large for this device's LDS (" + std::to_string(max_width) + " entries need " +
                       std::to_string(m->lds_bytes) + " bytes, at most " + std::to_string(max_lds) +
                       " per workgroup): use a smaller MC_OPT_ITEM_WINDOW");
    }

    // Persistent window kernel: 256-thread workgroups, windows staged through
    // registers (width <= 256 * kPersistPer), one workgroup per resident slot.
    m->persistent = m->persistent && m->block == 256 && max_width <= 256 * kPersistPer;
    // The M-step runs inside the E-step kernel when that kernel is the
    // iteration's only launch (every work item a window item, all of one
    // session format) in the segmented mode: the persistent kernel or the
    // one-workgroup-per-item kernel.
    m->fuse = m->fused_merge && m->reduce_mode == MC_REDUCE_SEGMENTED && m->n_win_items > 0 &&
              m->n_glob_items == 0 && m->n_dir_items == 0 &&
              (m->n_pk_items == 0 || m->n_pk_items == m->n_win_items);
    if (m->persistent && m->n_win_items > 0) {
        int n_cu = 0;
        MC_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, m->device));
        for (int pk = 0; pk < 2; pk++) {
            int per_cu = 0;
            MC_HIP(persistent_occupancy(m, pk != 0, &per_cu));
            // At most 6 workgroups per CU: the ten-rank PBM kernel (71 VGPRs)
            // fits 7, which measured 4 % slower (0.080 -> 0.084 ms per
            // iteration over 8M sessions) with one more window staged per CU.
            m->persist_grid[pk] = std::max(1, std::min(per_cu, 6) * n_cu);
        }
    }

    MC_TRY(dalloc(m, &m->d_packed, std::max<size_t>(h.packed.size() * 4, 16)));
    MC_TRY(dalloc(m, &m->d_pidx16, std::max<size_t>(h_pidx16.size() * 2, 4)));
    MC_TRY(dalloc(m, &m->d_pidx, std::max<size_t>(h_pidx.size() * 4, 4)));
    MC_TRY(dalloc(m, &m->d_mask, h_mask.size() * 2));
    if (m->varlen) MC_TRY(dalloc(m, &m->d_len, h_len.size()));
    MC_TRY(dalloc(m, &m->d_items, items.size() * sizeof(Item)));
    MC_TRY(dalloc(m, &m->d_shared_idx, shared_idx.size() * 4));
    MC_TRY(dalloc(m, &m->d_split_groups, std::max<size_t>(h.split_groups.size(), 1) * sizeof(int4)));
    MC_TRY(dalloc(m, &m->d_split_done, std::max<size_t>(h.split_groups.size(), 1) * 4));
    MC_TRY(dalloc(m, &m->d_grid_done, 4));
    MC_TRY(dalloc(m, &m->d_lists, lists.size() * 4));
    for (int t = 0; t < T; t++) {
        for (int b = 0; b < 2; b++) {
            MC_TRY(dalloc(m, &m->d_val[b][t], (size_t)n_qd * 4));
            MC_TRY(dalloc(m, &m->d_num[b][t], (size_t)n_qd * 4));
        }
        MC_TRY(dalloc(m, &m->d_gacc[t], (size_t)n_qd * 8));
    }
    for (int b = 0; b < 2; b++) {
        MC_TRY(dalloc(m, &m->d_sh_val[b], (size_t)K * 4));
        MC_TRY(dalloc(m, &m->d_sh_num[b], (size_t)K * 4));
        MC_TRY(dalloc(m, &m->d_sh_den[b], (size_t)K * 4));
    }
    MC_TRY(dalloc(m, &m->d_counts, (size_t)n_qd * 8));
    MC_TRY(dalloc(m, &m->d_counts16, (size_t)n_qd * 2));
    MC_TRY(dalloc(m, &m->d_sh_base, sh_base.size() * 8));
    MC_TRY(dalloc(m, &m->d_sh_acc_map, sh_map.size() * 4));
    MC_TRY(dalloc(m, &m->d_sh_acc, (size_t)KACC * 8));

    if (!h.packed.empty())
        MC_HIP(hipMemcpyAsync(m->d_packed, h.packed.data(), h.packed.size() * 4, hipMemcpyHostToDevice, m->stream));
    if (!h_pidx16.empty())
        MC_HIP(hipMemcpyAsync(m->d_pidx16, h_pidx16.data(), h_pidx16.size() * 2, hipMemcpyHostToDevice, m->stream));
    if (!h_pidx.empty())
        MC_HIP(hipMemcpyAsync(m->d_pidx, h_pidx.data(), h_pidx.size() * 4, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemcpyAsync(m->d_mask, h_mask.data(), h_mask.size() * 2, hipMemcpyHostToDevice, m->stream));
    if (m->varlen)
        MC_HIP(hipMemcpyAsync(m->d_len, h_len.data(), h_len.size(), hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemcpyAsync(m->d_items, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice, m->stream));
    if (!shared_idx.empty())
        MC_HIP(hipMemcpyAsync(m->d_shared_idx, shared_idx.data(), shared_idx.size() * 4, hipMemcpyHostToDevice,
                              m->stream));
    if (!h.split_groups.empty())
        MC_HIP(hipMemcpyAsync(m->d_split_groups, h.split_groups.data(), h.split_groups.size() * sizeof(int4),
                              hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemsetAsync(m->d_split_done, 0, std::max<size_t>(h.split_groups.size(), 1) * 4, m->stream));
    MC_HIP(hipMemsetAsync(m->d_grid_done, 0, 4, m->stream));
    if (!lists.empty())
        MC_HIP(hipMemcpyAsync(m->d_lists, lists.data(), lists.size() * 4, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemcpyAsync(m->d_counts, counts.data(), counts.size() * 8, hipMemcpyHostToDevice, m->stream));
    std::vector<uint16_t> counts16((size_t)n_qd);
    for (int64_t p = 0; p < n_qd; p++)
        counts16[(size_t)p] = counts[(size_t)p].x < 255 ? (uint16_t)(counts[(size_t)p].x | (counts[(size_t)p].y << 8))
                                                      : (uint16_t)kCountsEscape;
    MC_HIP(hipMemcpyAsync(m->d_counts16, counts16.data(), counts16.size() * 2, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemcpyAsync(m->d_sh_base, sh_base.data(), sh_base.size() * 8, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemcpyAsync(m->d_sh_acc_map, sh_map.data(), sh_map.size() * 4, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipMemsetAsync(m->d_sh_acc, 0, (size_t)KACC * 8, m->stream));
    for (int t = 0; t < T; t++) {
        MC_HIP(hipMemsetAsync(m->d_gacc[t], 0, (size_t)n_qd * 8, m->stream));
        for (int b = 0; b < 2; b++) {
            MC_TRY(fill(m, m->d_val[b][t], 0.5f, n_qd));
            MC_TRY(fill(m, m->d_num[b][t], 1.f, n_qd));
        }
        m->den_state[t] = 0;
    }
    for (int b = 0; b < 2; b++) {
        MC_TRY(fill(m, m->d_sh_val[b], 0.5f, K));
        MC_TRY(fill(m, m->d_sh_num[b], 1.f, K));
        MC_TRY(fill(m, m->d_sh_den[b], 2.f, K));
    }
    m->den_state[T] = 0;
    MC_HIP(hipStreamSynchronize(m->stream));
    m->cur = 0;
    m->estep_pending = false;
    m->initialized = true;
    return MC_OK;
}

mc_status mc_init_from_dataset(mc_model* m, const mc_dataset* ds) {
    MC_REQUIRE(m && ds, "null argument");
    const HostDataset& d = ds->d;
    return mc_init_parameters(m, (int64_t)d.train_query.size(), d.train_query.data(), d.train_pidx.data(),
                              d.train_clicks.data(), d.train_len.empty() ? nullptr : d.train_len.data(), d.n_qd);
}

mc_status mc_process_session(mc_model* m) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(!m->estep_pending, "E-step already pending; call mc_update_parameters");
    MC_HIP(hipSetDevice(m->device));
    MC_TRY(launch_estep(m, m->cur, 1));
    MC_HIP(hipStreamSynchronize(m->stream));
    m->estep_pending = true;
    return MC_OK;
}

mc_status mc_update_parameters(mc_model* m) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(m->estep_pending, "no pending E-step; call mc_process_session first");
    MC_HIP(hipSetDevice(m->device));
    MC_TRY(launch_mstep(m, m->cur, 1));
    MC_HIP(hipStreamSynchronize(m->stream));
    m->cur ^= 1;
    m->estep_pending = false;
    for (int t = 0; t <= m->T; t++) m->den_state[t] = 2;
    return MC_OK;
}

mc_status mc_em_iterations(mc_model* m, int n_iter, int use_graph) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(!m->estep_pending, "an E-step is pending");
    MC_REQUIRE(n_iter >= 0, "n_iter must be >= 0");
    if (n_iter == 0) return MC_OK;
    MC_HIP(hipSetDevice(m->device));
    if (use_graph) {
        hipGraphExec_t g = nullptr;
        MC_TRY(get_graph(m, n_iter, m->cur, &g));
        MC_HIP(hipGraphLaunch(g, m->stream));
    } else {
        int c = m->cur;
        for (int i = 0; i < n_iter; i++) {
            const int wn = i == n_iter - 1;
            MC_TRY(launch_estep(m, c, wn));
            MC_TRY(launch_mstep(m, c, wn));
            c ^= 1;
        }
    }
    MC_HIP(hipStreamSynchronize(m->stream));
    m->cur ^= (n_iter & 1);
    for (int t = 0; t <= m->T; t++) m->den_state[t] = 2;
    return MC_OK;
}

mc_status mc_em_iterations_timed(mc_model* m, int n_iter, double* ms_total, double* ms_estep_avg,
                                 double* ms_mstep_avg) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(!m->estep_pending, "an E-step is pending");
    MC_REQUIRE(n_iter >= 1, "n_iter must be >= 1");
    MC_HIP(hipSetDevice(m->device));
    std::vector<hipEvent_t> ev((size_t)(3 * n_iter + 1));
    for (auto& e : ev) MC_HIP(hipEventCreate(&e));
    int c = m->cur;
    MC_HIP(hipEventRecord(ev[0], m->stream));
    for (int i = 0; i < n_iter; i++) {
        const int wn = i == n_iter - 1;
        MC_HIP(hipEventRecord(ev[3 * i + 1], m->stream));
        MC_TRY(launch_estep(m, c, wn));
        MC_HIP(hipEventRecord(ev[3 * i + 2], m->stream));
        MC_TRY(launch_mstep(m, c, wn));
        MC_HIP(hipEventRecord(ev[3 * i + 3], m->stream));
        c ^= 1;
    }
    MC_HIP(hipStreamSynchronize(m->stream));
    double te = 0, tm = 0;
    for (int i = 0; i < n_iter; i++) {
        float a = 0, b = 0;
        MC_HIP(hipEventElapsedTime(&a, ev[3 * i + 1], ev[3 * i + 2]));
        MC_HIP(hipEventElapsedTime(&b, ev[3 * i + 2], ev[3 * i + 3]));
        te += a;
        tm += b;
    }
    float tot = 0;
    MC_HIP(hipEventElapsedTime(&tot, ev[0], ev[3 * n_iter]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (ms_total) *ms_total = tot;
    if (ms_estep_avg) *ms_estep_avg = te / n_iter;
    if (ms_mstep_avg) *ms_mstep_avg = tm / n_iter;
    m->cur ^= (n_iter & 1);
    for (int t = 0; t <= m->T; t++) m->den_state[t] = 2;
    return MC_OK;
}

mc_status mc_em_graph_timed(mc_model* m, int n_iter, double* ms_total) {
    MC_REQUIRE(m && m->initialized && n_iter >= 1, "invalid argument");
    MC_REQUIRE(!m->estep_pending, "an E-step is pending");
    MC_HIP(hipSetDevice(m->device));
    // One warm replay, then the timed replay (graphs built outside the timing).
    MC_TRY(mc_em_iterations(m, n_iter, 1));
    hipGraphExec_t g = nullptr;
    MC_TRY(get_graph(m, n_iter, m->cur, &g));
    hipEvent_t a, b;
    MC_HIP(hipEventCreate(&a));
    MC_HIP(hipEventCreate(&b));
    MC_HIP(hipEventRecord(a, m->stream));
    MC_HIP(hipGraphLaunch(g, m->stream));
    MC_HIP(hipEventRecord(b, m->stream));
    MC_HIP(hipStreamSynchronize(m->stream));
    float ms = 0;
    MC_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    m->cur ^= (n_iter & 1);
    for (int t = 0; t <= m->T; t++) m->den_state[t] = 2;
    if (ms_total) *ms_total = ms;
    return MC_OK;
}

mc_status mc_synchronize(mc_model* m) {
    MC_REQUIRE(m, "null model");
    MC_HIP(hipStreamSynchronize(m->stream));
    return MC_OK;
}

mc_status mc_get_n_tables(const mc_model* m, int* n_tables) {
    MC_REQUIRE(m && n_tables, "null argument");
    *n_tables = m->T + 1;
    return MC_OK;
}

mc_status mc_get_table_info(const mc_model* m, int t, const char** name, int* is_public, int64_t* size) {
    MC_REQUIRE(m && t >= 0 && t <= m->T, "bad table index");
    if (name) *name = table_name(m->type, t);
    if (is_public) *is_public = t == m->T;
    if (size) *size = t == m->T ? m->K : m->n_qd;
    return MC_OK;
}

mc_status mc_get_parameters(mc_model* m, int t, float* num, float* den) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(t >= 0 && t <= m->T, "bad table index");
    MC_HIP(hipSetDevice(m->device));
    MC_HIP(hipStreamSynchronize(m->stream));
    const bool shared = t == m->T;
    const int64_t n = shared ? m->K : m->n_qd;
    if (!shared && m->den_state[t] == 3) {
        // Values set by the gradient path: (value, 1).
        if (num) MC_HIP(hipMemcpy(num, m->d_val[m->cur][t], n * 4, hipMemcpyDeviceToHost));
        if (den) std::fill(den, den + n, 1.f);
        return MC_OK;
    }
    if (num) MC_HIP(hipMemcpy(num, shared ? m->d_sh_num[m->cur] : m->d_num[m->cur][t], n * 4, hipMemcpyDeviceToHost));
    if (den) {
        if (shared) {
            MC_HIP(hipMemcpy(den, m->d_sh_den[m->cur], n * 4, hipMemcpyDeviceToHost));
        } else if (m->den_state[t] == 0) {
            std::fill(den, den + n, 2.f);
        } else if (m->den_state[t] == 1) {
            MC_HIP(hipMemcpy(den, m->d_den_user[t], n * 4, hipMemcpyDeviceToHost));
        } else {
            // Same arithmetic as base_den() in the kernels.
            for (int64_t i = 0; i < n; i++) {
                const float o = (float)m->h_counts[i].x, c = (float)m->h_counts[i].y;
                if (m->type == MC_CCM) den[i] = 2.f + o + c;
                else if (m->type == MC_DBN && t == 1) den[i] = 2.f + 2.f * o - c;
                else den[i] = 2.f + o;
            }
        }
    }
    return MC_OK;
}

mc_status mc_set_parameters(mc_model* m, int t, const float* num, const float* den) {
    MC_REQUIRE(m && m->initialized && num && den, "invalid argument");
    MC_REQUIRE(t >= 0 && t <= m->T, "bad table index");
    MC_REQUIRE(!m->estep_pending, "an E-step is pending");
    MC_HIP(hipSetDevice(m->device));
    const bool shared = t == m->T;
    const int64_t n = shared ? m->K : m->n_qd;
    std::vector<float> val((size_t)n);
    for (int64_t i = 0; i < n; i++) {   // Param::value, param.cu:75-87
        const float v = num[i] / den[i];
        val[i] = v < kClampMax ? v : kClampMax;
    }
    float* dv = shared ? m->d_sh_val[m->cur] : m->d_val[m->cur][t];
    float* dn = shared ? m->d_sh_num[m->cur] : m->d_num[m->cur][t];
    MC_HIP(hipMemcpy(dv, val.data(), n * 4, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(dn, num, n * 4, hipMemcpyHostToDevice));
    if (shared) {
        MC_HIP(hipMemcpy(m->d_sh_den[m->cur], den, n * 4, hipMemcpyHostToDevice));
    } else {
        if (!m->d_den_user[t]) MC_TRY(dalloc(m, &m->d_den_user[t], (size_t)n * 4));
        MC_HIP(hipMemcpy(m->d_den_user[t], den, n * 4, hipMemcpyHostToDevice));
        m->den_state[t] = 1;
    }
    return MC_OK;
}

mc_status mc_reset_parameters(mc_model* m) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_HIP(hipSetDevice(m->device));
    for (int t = 0; t < m->T; t++) {
        MC_TRY(fill(m, m->d_val[m->cur][t], 0.5f, m->n_qd));
        MC_TRY(fill(m, m->d_num[m->cur][t], 1.f, m->n_qd));
        m->den_state[t] = 0;
    }
    MC_TRY(fill(m, m->d_sh_val[m->cur], 0.5f, m->K));
    MC_TRY(fill(m, m->d_sh_num[m->cur], 1.f, m->K));
    MC_TRY(fill(m, m->d_sh_den[m->cur], 2.f, m->K));
    MC_HIP(hipStreamSynchronize(m->stream));
    m->estep_pending = false;
    return MC_OK;
}

mc_status mc_get_memory_usage(const mc_model* m, size_t* bytes) {
    MC_REQUIRE(m && bytes, "null argument");
    *bytes = m->mem_bytes;
    return MC_OK;
}

mc_status mc_compute_memory_footprint(int model_type, int64_t n, int64_t n_qd, size_t* bytes) {
    MC_REQUIRE(bytes && model_type >= 0 && model_type <= 3 && n >= 0 && n_qd >= 0, "invalid argument");
    int T, K, KACC;
    shape_of(model_type, &T, &K, &KACC);
    // EM buffers, for the default layout.  Sessions: at most 40 B of indices
    // (20 B for sessions of window items), 2 B mask, 1 B length.  Per entry
    // and table: two values, two numerators, an 8-byte accumulator; per
    // entry: 8 B counts, 2 B packed counts, 4 B if merged across items.
    // Work items of at least 64 sessions (24 B each, 8 B of batch lists).
    const size_t items = (size_t)(n / 64 + 1);
    // Shared tables: values, numerators, denominators (two each), FP64 bases,
    // accumulator map; plus the 16-byte minimum of small allocations.
    size_t b = (size_t)n * 43 + (size_t)n_qd * ((size_t)T * 24 + 14) + (size_t)K * 64;
    b += (size_t)KACC * 8 + items * (sizeof(Item) + 8) + 256;
    *bytes = b;
    return MC_OK;
}

mc_status mc_get_layout_info(const mc_model* m, int64_t* n_items, int64_t* n_shared_entries,
                             int64_t* max_item_sessions, int64_t* n_em_sessions) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    if (n_items) *n_items = m->n_items;
    if (n_shared_entries) *n_shared_entries = m->n_shared;
    if (max_item_sessions) *max_item_sessions = m->max_item_sessions;
    if (n_em_sessions) *n_em_sessions = m->n_em_sessions;
    return MC_OK;
}

mc_status mc_get_stream_info(const mc_model* m, int64_t* n_window_sessions, int64_t* index_bytes) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    if (n_window_sessions) *n_window_sessions = m->n_w;
    if (index_bytes) *index_bytes = m->n_pk * 16 + (m->n_w - m->n_pk) * kMaxSerp * 2 + (m->N - m->n_w) * kMaxSerp * 4;
    return MC_OK;
}

mc_status mc_get_session_formats(const mc_model* m, int64_t* n_packed, int64_t* n_window16, int64_t* n_absolute) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    if (n_packed) *n_packed = m->n_pk;
    if (n_window16) *n_window16 = m->n_w - m->n_pk;
    if (n_absolute) *n_absolute = m->N - m->n_w;
    return MC_OK;
}

void mc_model_destroy(mc_model* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    (void)hipStreamSynchronize(m->stream);
    release(m);
    (void)hipStreamDestroy(m->stream);
    delete m;
}

}  // extern "C"
