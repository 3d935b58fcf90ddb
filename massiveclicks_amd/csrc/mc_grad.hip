// mc_grad.hip -- log-likelihood gradient kernels and the Adam fit.
//
// The objective is the click log-likelihood the reference evaluates
// (get_log_conditional_click_probs, e.g. dbn.cu:611-639, summed over the
// ranks of every training session), as a function of logit-parameterised
// tables: value = sigmoid(logit) for attractiveness, examination,
// satisfaction and continuation.  One lane runs one session: the forward
// chain is recomputed in registers and differentiated in reverse, so no
// per-session state is stored.  Gradients are summed per query-document
// entry with the same work items, LDS windows and order-independent 64-bit
// fixed-point sums as the EM E-step (two's complement: contributions may be
// negative), so they are deterministic.
//
// The Adam fit maximises log-likelihood + log-prior, where the prior of each
// parameter is Beta(1 + w, 1 + w) with w the pseudo-count weight that makes
// plain EM's update a MAP-EM step: w = 1 for every parameter, plus the
// never-written temporaries the reference counts as (1,2) (DBN satisfaction:
// unclicked occurrences; PBM/UBM examination: sessions not writing the slot).
// Minibatches are runs of whole query groups, so an entry's gradient is
// complete whenever its batch runs; the shared (rank-indexed) tables get the
// batch gradient scaled to the whole training set.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"
#include "mc_kernels.h"
#include "mc_model.h"

namespace mc {
namespace {

constexpr float kLogitMax = 13.8f;   // sigmoid(13.8) = 1 - 1.0e-6, the Param::value clamp
constexpr int kShAccStride = 64;     // shared-table entries (>= 55 accumulators)
// A step's shared-table sums are split over kShParts copies, one per XCD
// (consecutive workgroups run on consecutive XCDs), so that the items'
// atomics do not all queue on one address; readers add the copies.
constexpr int kShParts = 8;
constexpr int kShAccSlot = kShParts * kShAccStride;

struct GArgs {
    const uint32_t* packed;
    const uint16_t* pidx16;
    const int32_t* pidx;
    const uint16_t* mask;
    const uint8_t* len;
    int64_t N, n_w, n_pk;
    const Item* items;
    const int32_t* lists;
    float* val[2];                 // per-qd values (read; written by Adam)
    float* logit[2];
    float* am[2];
    float* av[2];
    float* grad[2];                // gradient output (GRAD mode)
    unsigned long long* gacc[2];   // fixed-point sums of split entries
    const int2* counts;
    const uint16_t* counts16;     // packed counts (unpack_counts)
    float* sh_val;
    float* sh_num;
    float* sh_den;
    float* sh_logit;               // Adam state of the shared tables: the version read
    float* sh_am;
    float* sh_av;
    float* wr_logit;               // ... and the version written (may be the same)
    float* wr_am;
    float* wr_av;
    float* wr_val;
    int32_t wr_main;               // the written version is the model's own (also set num/den)
    float* sh_grad;
    const double* sh_prior;
    unsigned long long* sh_acc;
    double* ll_part;               // per item
    const int32_t* shared_idx;
    int32_t* steps;                // [0]: completed epochs (every batch once)
    int64_t shared_off, shared_cnt;  // split entries of the batch
    // Fused finalisation of the previous step (Adam epochs replayed from a
    // graph): its batch (-1: none), split entries, shared-table sums, scale.
    int32_t prev_batch;
    int64_t prev_off, prev_cnt;
    double prev_scale;
    const unsigned long long* prev_acc;
    unsigned long long* zero_acc;  // shared sums cleared by workgroup 0 (read two steps ago)
    int32_t K, kacc, max_width, batch, n_batches, use_prior;
    float lr, b1, b2, eps;
    double sh_scale;               // training sessions / batch sessions
};

// Signed 32.32 fixed point (two's complement), saturating at +-2^30.
__device__ __forceinline__ long long to_fix_signed(float x) {
    x = fminf(fmaxf(x, -1073741824.f), 1073741824.f);
    const float fl = floorf(x);
    const int hi = (int)fl;
    const uint32_t lo = (uint32_t)((x - fl) * kFixScale);
    return (long long)(((unsigned long long)(long long)hi << 32) | lo);
}

__device__ __forceinline__ double fix_to_double(unsigned long long v) { return (double)(long long)v * kFixInv; }

template <int T>
struct EmitLdsSigned {
    unsigned long long* acc;
    __device__ __forceinline__ void add(int t, int, int p, float x) {
        atomicAdd(&acc[p * T + t], (unsigned long long)to_fix_signed(x));
    }
};
struct EmitGlobalSigned {
    unsigned long long* g0;
    unsigned long long* g1;
    __device__ __forceinline__ void add(int t, int, int p, float x) {
        atomicAdd(t == 0 ? &g0[p] : &g1[p], (unsigned long long)to_fix_signed(x));
    }
};

__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + __expf(-x)); }

// ---------------------------------------------------------------------------
// Per-session gradients (logit space).  Clicked occurrences of an
// attractiveness entry contribute exactly 1 - a each (d log a / d logit);
// that term is added per entry from its click count when the entry is
// finalised, so only the other contributions are emitted.  acc[k] collects
// the gradient of shared-table entry k, ll the session's log-likelihood.
// ---------------------------------------------------------------------------

// PBM/UBM: log P(C) = sum_r c log(a e) + (1 - c) log(1 - a e), e from
// rank r (PBM) or slot r(r+1)/2 + prev (UBM, ubm.cu:201).
template <int MODEL, bool VARLEN, bool LL, class Src, class Emit>
__device__ __forceinline__ void grad_pbm_ubm(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                             const float* sh, float (&acc)[55], float& ll) {
    int slot[kL];
    if (MODEL == kUBM) ubm_slots(mm, slot);
#pragma unroll
    for (int r = 0; r < kL; r++) {
        if (VARLEN && r >= L) continue;
        const int p = pv[r];
        const float a = src.get(0, p);
        const float e = MODEL == kUBM ? sh[slot[r]] : sh[r];
        const float ae = a * e;
        float ge;
        if ((mm >> r) & 1u) {
            if (LL) ll += __logf(ae);
            ge = 1.f - e;
        } else {
            const float q = 1.f - ae;
            if (LL) ll += __logf(q);
            const float w = ae * __builtin_amdgcn_rcpf(q);
            em.add(0, r, p, -w * (1.f - a));
            ge = -w * (1.f - e);
        }
        if (MODEL == kUBM) {
            const int row = r * (r + 1) / 2;
#pragma unroll
            for (int j = 0; j < r; j++) acc[row + j] += (slot[r] == row + j) ? ge : 0.f;
            acc[row + kL - 1] += (slot[r] == row + kL - 1) ? ge : 0.f;
        } else {
            acc[r] += ge;
        }
    }
}

// DBN / CCM: the conditional chain of dbn.cu:611-639 / ccm.cu:604-630.
// Forward: ex_0 = 1; click: p = a ex, ex' = g(1-s) [CCM: t2(1-a) + t3 a];
// no click: p = 1 - a ex, ex' = ex g(1-a)/p [CCM: t1].  Reverse sweep with
// dex = d LL / d ex_{r+1}:
//   click     d/da = 1/a (+ dex (t3 - t2) for the CCM), d/ds = -g dex,
//             d/dg += dex (1-s), d LL/d ex_r = 1/ex_r;
//   no click  q = 1 - a ex: d/da = -ex/q + dex ex g (ex-1)/q^2,
//             d/dg += dex ex (1-a)/q, d LL/d ex_r = -a/q + dex g (1-a)/q^2.
template <int MODEL, bool VARLEN, bool LL, class Src, class Emit>
__device__ __forceinline__ void grad_cascade(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                             const float* sh, float (&acc)[55], float& ll) {
#pragma clang fp contract(fast)
    const float g = sh[0];                                 // DBN gamma / CCM tau1
    const float t2 = MODEL == kCCM ? sh[1] : 0.f, t3 = MODEL == kCCM ? sh[2] : 0.f;
    float exv[kL];
    float e = 1.f;
#pragma unroll
    for (int r = 0; r < kL; r++) {
        exv[r] = e;
        if (VARLEN && r >= L) continue;
        const float a = src.get(0, pv[r]);
        if ((mm >> r) & 1u) {
            if (LL) ll += __logf(a * e);
            e = MODEL == kDBN ? g * (1.f - src.get(1, pv[r])) : t2 * (1.f - a) + t3 * a;
        } else {
            const float q = 1.f - a * e;
            if (LL) ll += __logf(q);
            e = e * g * (1.f - a) * __builtin_amdgcn_rcpf(q);
        }
    }
    __asm__ volatile("" ::: "memory");
    float dex = 0.f, dg = 0.f, d2 = 0.f, d3 = 0.f;
#pragma unroll
    for (int r = kL - 1; r >= 0; r--) {
        if (VARLEN && r >= L) continue;
        const float E = exv[r];
        const int p = pv[r];
        const float a = src.get(0, p);
        if ((mm >> r) & 1u) {
            if (MODEL == kDBN) {
                const float s = src.get(1, p);
                em.add(1, r, p, -dex * g * s * (1.f - s));
                dg += dex * (1.f - s);
            } else {
                em.add(0, r, p, dex * (t3 - t2) * a * (1.f - a));
                d2 += dex * (1.f - a);
                d3 += dex * a;
            }
            dex = __builtin_amdgcn_rcpf(E);
        } else {
            const float iq = __builtin_amdgcn_rcpf(1.f - a * E);
            const float iq2 = iq * iq;
            const float da = -E * iq + dex * E * g * (E - 1.f) * iq2;
            dg += dex * E * (1.f - a) * iq;
            dex = -a * iq + dex * g * (1.f - a) * iq2;
            em.add(0, r, p, da * a * (1.f - a));
        }
    }
    acc[0] += dg * g * (1.f - g);
    if (MODEL == kCCM) {
        acc[1] += d2 * t2 * (1.f - t2);
        acc[2] += d3 * t3 * (1.f - t3);
    }
}

template <int MODEL, bool VARLEN, bool LL, class Src, class Emit>
__device__ __forceinline__ void grad_session(const int (&pv)[kL], int L, uint32_t m, const Src& src, Emit& em,
                                             const float* sh, float (&acc)[55], float& ll) {
    const uint32_t mm = m & ((1u << L) - 1u);
    if (MODEL == kPBM || MODEL == kUBM) grad_pbm_ubm<MODEL, VARLEN, LL>(pv, L, mm, src, em, sh, acc, ll);
    else grad_cascade<MODEL, VARLEN, LL>(pv, L, mm, src, em, sh, acc, ll);
}

// Adam (maximising): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// logit += lr/(1-b1^t) m / (sqrt(v/(1-b2^t)) + eps), clamped so that the
// value stays inside Param::value's range.
struct AdamStep {
    float lr_bc1, inv_sqrt_bc2, b1, b2, eps;
    __device__ __forceinline__ float apply(float& logit, float& m, float& v, float g) const {
        m = b1 * m + (1.f - b1) * g;
        v = b2 * v + (1.f - b2) * g * g;
        const float d = sqrtf(v) * inv_sqrt_bc2 + eps;
        logit = fminf(fmaxf(logit + lr_bc1 * m / d, -kLogitMax), kLogitMax);
        return sigmoid(logit);
    }
};

__device__ __forceinline__ AdamStep adam_step(const GArgs& G, int t_step) {
    AdamStep s;
    const float bc1 = 1.f - powf(G.b1, (float)t_step);
    const float bc2 = 1.f - powf(G.b2, (float)t_step);
    s.lr_bc1 = G.lr / bc1;
    s.inv_sqrt_bc2 = 1.f / sqrtf(bc2);
    s.b1 = G.b1;
    s.b2 = G.b2;
    s.eps = G.eps;
    return s;
}

// Gradient of per-qd entry p of table t from its fixed-point sum: adds the
// clicked occurrences' 1 - a (attractiveness) and, for Adam, the prior.
template <int MODEL>
__device__ __forceinline__ float entry_gradient(int t, double sum, int2 oc, float val, bool prior) {
    double g = sum;
    if (t == 0) g += (double)oc.y * (1.0 - (double)val);
    if (prior) {
        const double w = 1.0 + ((MODEL == kDBN && t == 1) ? (double)(oc.x - oc.y) : 0.0);
        g += w * (1.0 - 2.0 * (double)val);
    }
    return (float)g;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Merge work of grad_merge_kernel.
__device__ __forceinline__ unsigned long long load_sum(const unsigned long long* p) { return *p; }
// The sum of the kShParts copies of shared-table entry p[0].
__device__ __forceinline__ unsigned long long shared_sum(const unsigned long long* p) {
    unsigned long long s = 0;
#pragma unroll
    for (int q = 0; q < kShParts; q++) s += p[q * kShAccStride];
    return s;
}

// Adam step of shared-table entry k from the sum of a step of batch
// `batch` (scaled by `scale`); returns the new value and, when `write`,
// stores the new state in the written version.
__device__ __forceinline__ float shared_adam(const GArgs& G, int k, double sum, double scale, int batch, bool write) {
    const float val = G.sh_val[k];
    double g = sum * scale;
    if (G.use_prior) g += G.sh_prior[k] * (1.0 - 2.0 * (double)val);
    const AdamStep ad = adam_step(G, G.steps[0] * G.n_batches + batch + 1);
    float logit = G.sh_logit[k], am = G.sh_am[k], av = G.sh_av[k];
    const float nv = ad.apply(logit, am, av, (float)g);
    if (write) {
        G.wr_logit[k] = logit;
        G.wr_am[k] = am;
        G.wr_av[k] = av;
        G.wr_val[k] = nv;
        if (G.wr_main) {
            G.sh_num[k] = nv;
            G.sh_den[k] = 1.f;
        }
    }
    return nv;
}

template <int MODEL, bool ADAM>
__device__ __forceinline__ void merge_shared_entry(const GArgs& G, int k) {
    const double sum = fix_to_double(shared_sum(&G.sh_acc[k]));
    if (ADAM) shared_adam(G, k, sum, G.sh_scale, G.batch, true);
    else G.sh_grad[k] = (float)sum;
}

template <int MODEL, bool ADAM>
__device__ __forceinline__ void merge_split_entry(const GArgs& G, int64_t off, int64_t i, const AdamStep& ad) {
    constexpr int T = Shape<MODEL>::T;
    const int p = G.shared_idx[off + i];
    const int2 oc = G.counts[p];
#pragma unroll
    for (int t = 0; t < T; t++) {
        const float val = G.val[t][p];
        const float g = entry_gradient<MODEL>(t, fix_to_double(load_sum(&G.gacc[t][p])), oc, val, ADAM && G.use_prior);
        if (ADAM) G.val[t][p] = ad.apply(G.logit[t][p], G.am[t][p], G.av[t][p], g);
        else G.grad[t][p] = g;
        G.gacc[t][p] = 0ull;
    }
}

// One workgroup per listed item: the sessions' gradients, reduced in the
// item's LDS window (WIN) or with global fixed-point atomics, then
// finalised for the entries the item owns: written out (ADAM = false) or
// applied by Adam in place.
template <int MODEL, bool VARLEN, bool WIN, bool ADAM, int NT>
__global__ void __launch_bounds__(NT) grad_item_kernel(GArgs G, int64_t list_off) {
    constexpr int T = Shape<MODEL>::T;
    constexpr int K = Shape<MODEL>::K;
    constexpr int NW = NT / 64;
    extern __shared__ unsigned long long lds[];
    __shared__ float sh[K < 64 ? 64 : K];
    __shared__ long long red[K][NW];
    __shared__ double lred[NW];

    const int item_idx = G.lists[list_off + blockIdx.x];
    const Item it = G.items[item_idx];
    const int tid = threadIdx.x;
    const int W = WIN ? it.width : 0;
    unsigned long long* lacc = lds;
    float* lval = (float*)(lds + T * G.max_width);

    KArgs ka;
    ka.packed = G.packed;
    ka.pidx16 = G.pidx16;
    ka.pidx = G.pidx;
    ka.mask = G.mask;
    ka.len = G.len;
    ka.N = G.N;
    ka.n_w = G.n_w;
    ka.n_pk = G.n_pk;
    const Stream st = make_stream(ka);
    // Window items are packed (16-byte session records) or rank-major.
    const bool packed = WIN && (it.flags & kItemPacked) != 0;
    uint32_t m0 = 0;
    int L0 = 0;
    int pv0[kL];
    PackedSession r0;
    if (packed) r0 = load_packed(st, it.s0 + tid, it.s1);
    else load_session<VARLEN, WIN>(st, it.s0 + tid, it.s1, m0, L0, pv0);
    for (int k = tid; k < K; k += NT) sh[k] = G.sh_val[k];
    if (ADAM && G.prev_batch >= 0) {
        // The previous step's finalisation: every workgroup computes the
        // shared tables' new values for its own sessions (workgroup 0 stores
        // them), and the workgroups share the previous batch's split entries.
        for (int k = tid; k < K; k += NT)
            sh[k] = shared_adam(G, k, fix_to_double(shared_sum(&G.prev_acc[k])), G.prev_scale, G.prev_batch, blockIdx.x == 0);
        if (blockIdx.x == 0)
            for (int j = tid; j < kShAccSlot; j += NT) G.zero_acc[j] = 0ull;
        const AdamStep pad = adam_step(G, G.steps[0] + 1);
        for (int64_t i = (int64_t)blockIdx.x * NT + tid; i < G.prev_cnt; i += (int64_t)gridDim.x * NT)
            merge_split_entry<MODEL, true>(G, G.prev_off, i, pad);
    }
    if (WIN) {
        for (int i = tid; i < W; i += NT) {
#pragma unroll
            for (int t = 0; t < T; t++) {
                lacc[i * T + t] = 0ull;
                lval[i * T + t] = G.val[t][it.base + i];
            }
        }
    }
    __syncthreads();

    // Wide workgroups (small minibatches, one round of items) fetch the
    // inputs of their entries' finalisation now, so that they arrive during
    // the session loop: lane tid owns entries tid + u * NT, u < PER.
    constexpr int PER = NT >= 1024 ? 2 : 0;
    constexpr int PERA = PER > 0 ? PER : 1;
    const bool interior = WIN && (it.flags & kItemInterior) != 0;
    uint32_t f_oc[PERA];
    float f_logit[PERA][T], f_m[PERA][T], f_v[PERA][T];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = u * NT + tid;
        const bool on = interior && i < W;
        f_oc[u] = on ? G.counts16[it.base + i] : 0u;
#pragma unroll
        for (int t = 0; t < T; t++) {
            f_logit[u][t] = (ADAM && on) ? G.logit[t][it.base + i] : 0.f;
            f_m[u][t] = (ADAM && on) ? G.am[t][it.base + i] : 0.f;
            f_v[u][t] = (ADAM && on) ? G.av[t][it.base + i] : 0.f;
        }
    }

    float acc[55];
#pragma unroll
    for (int k = 0; k < 55; k++) acc[k] = 0.f;
    float ll = 0.f;
    {
        const SrcLds<T> lsrc{lval};
        const SrcGlobal gsrc{G.val[0], G.val[1]};
        EmitLdsSigned<T> el{lacc};
        EmitGlobalSigned eg{G.gacc[0], G.gacc[1]};
        if (packed) {
            for (int s = it.s0 + tid; s < it.s1; s += NT) {
                const PackedSession nxt = load_packed(st, s + NT, it.s1);
                decode_packed<VARLEN>(r0, m0, L0, pv0);
                grad_session<MODEL, VARLEN, !ADAM>(pv0, L0, m0, lsrc, el, sh, acc, ll);
                r0 = nxt;
            }
        }
        for (int s = packed ? it.s1 : it.s0 + tid; s < it.s1; s += NT) {
            uint32_t m2;
            int L2;
            int pv2[kL];
            load_session<VARLEN, WIN>(st, s + NT, it.s1, m2, L2, pv2);
            if (WIN) grad_session<MODEL, VARLEN, !ADAM>(pv0, L0, m0, lsrc, el, sh, acc, ll);
            else grad_session<MODEL, VARLEN, !ADAM>(pv0, L0, m0, gsrc, eg, sh, acc, ll);
            m0 = m2;
            L0 = L2;
#pragma unroll
            for (int r = 0; r < kL; r++) pv0[r] = pv2[r];
        }
    }

    // Shared tables and the log-likelihood: wave sums, then one atomic per
    // accumulator and item.
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const long long v = wave_sum(to_fix_signed(acc[k]));
        if (lane == 0) red[k][wave] = v;
    }
    if (!ADAM) {
        const double lw = wave_sum_f64((double)ll);
        if (lane == 0) lred[wave] = lw;
    }
    __syncthreads();
    for (int k = tid; k < K; k += NT) {
        long long v = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) v += red[k][w];
        if (v != 0) atomicAdd(&G.sh_acc[(blockIdx.x % kShParts) * kShAccStride + k], (unsigned long long)v);
    }
    if (!ADAM && tid == 0) {
        double lsum = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) lsum += lred[w];
        G.ll_part[item_idx] = lsum;
    }

    if (WIN) {
        AdamStep ad;
        if (ADAM && interior) ad = adam_step(G, G.steps[0] + 1);
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = u * NT + tid;
            if (!interior || i >= W) continue;
            const int p = it.base + i;
            const int2 oc = unpack_counts(f_oc[u], G.counts, p);
#pragma unroll
            for (int t = 0; t < T; t++) {
                const float g = entry_gradient<MODEL>(t, fix_to_double(lacc[i * T + t]), oc, lval[i * T + t],
                                                      ADAM && G.use_prior);
                if (ADAM) {
                    G.val[t][p] = ad.apply(f_logit[u][t], f_m[u][t], f_v[u][t], g);
                    G.logit[t][p] = f_logit[u][t];
                    G.am[t][p] = f_m[u][t];
                    G.av[t][p] = f_v[u][t];
                } else {
                    G.grad[t][p] = g;
                }
            }
        }
        // The remaining entries, four per lane and pass, with all of a pass's
        // loads issued before its first store.
        constexpr int U = 4;
        for (int i0 = tid + (interior ? PER * NT : 0); i0 < W; i0 += U * NT) {
            uint32_t oc[U];
            float lg[U][T], am[U][T], av[U][T];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT;
                const bool on = interior && i < W;
                const int p = it.base + (on ? i : 0);
                oc[u] = on ? G.counts16[p] : 0u;
#pragma unroll
                for (int t = 0; t < T; t++) {
                    lg[u][t] = (ADAM && on) ? G.logit[t][p] : 0.f;
                    am[u][t] = (ADAM && on) ? G.am[t][p] : 0.f;
                    av[u][t] = (ADAM && on) ? G.av[t][p] : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT;
                if (i >= W) continue;
                const int p = it.base + i;
                const int2 ocu = interior ? unpack_counts(oc[u], G.counts, p) : make_int2(0, 0);
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const unsigned long long v = lacc[i * T + t];
                    if (interior) {
                        const float g = entry_gradient<MODEL>(t, fix_to_double(v), ocu, lval[i * T + t],
                                                              ADAM && G.use_prior);
                        if (ADAM) {
                            G.val[t][p] = ad.apply(lg[u][t], am[u][t], av[u][t], g);
                            G.logit[t][p] = lg[u][t];
                            G.am[t][p] = am[u][t];
                            G.av[t][p] = av[u][t];
                        } else {
                            G.grad[t][p] = g;
                        }
                    } else if (v != 0ull) {
                        atomicAdd(&G.gacc[t][p], v);
                    }
                }
            }
        }
    }
}

// Block 0: the shared tables (gradient scaled to the training set, prior,
// Adam); other blocks: the listed split entries.  With Adam, the last block
// to finish advances the step counts, after every block has read them.
template <int MODEL, bool ADAM>
__global__ void __launch_bounds__(256) grad_merge_kernel(GArgs G) {
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        if (tid < G.K) merge_shared_entry<MODEL, ADAM>(G, tid);
        __syncthreads();
        for (int j = tid; j < kShAccSlot; j += blockDim.x) G.sh_acc[j] = 0ull;
        if (G.zero_acc)
            for (int j = tid; j < 3 * kShAccSlot; j += blockDim.x) G.zero_acc[j] = 0ull;
    } else {
        const int64_t i = (int64_t)(blockIdx.x - 1) * blockDim.x + tid;
        AdamStep ad;
        if (ADAM) ad = adam_step(G, G.steps[0] + 1);
        if (i < G.shared_cnt) merge_split_entry<MODEL, ADAM>(G, G.shared_off, i, ad);
    }
}

// Counts a completed epoch; launched after the step of the last batch.  A
// step of batch b in epoch e is Adam step e + 1 of b's entries and step
// e * batches + b + 1 of the shared tables.
__global__ void advance_epoch_kernel(int32_t* steps) {
    if (threadIdx.x == 0) steps[0] += 1;
}

// logit = log(v / (1 - v)) of every entry; moments to zero.
__global__ void logit_init_kernel(const float* val, float* logit, float* am, float* av, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = val[i];
    logit[i] = fminf(fmaxf(logf(v) - logf(1.f - v), -kLogitMax), kLogitMax);
    am[i] = 0.f;
    av[i] = 0.f;
}

__global__ void sigmoid_kernel(const float* logit, float* val, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) val[i] = sigmoid(logit[i]);
}

__global__ void zero_steps_kernel(int32_t* steps, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) steps[i] = 0;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

mc_status check_grad_support(mc_model* m) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    if (m->reduce_mode != MC_REDUCE_SEGMENTED || m->n_dir_items > 0) {
        set_error("the gradient path needs the segmented layout without direct items");
        return MC_ERR_UNSUPPORTED;
    }
    MC_REQUIRE(!m->estep_pending, "an E-step is pending");
    return MC_OK;
}

mc_status ensure_grad_buffers(mc_model* m) {
    if (m->grad_ready) return MC_OK;
    const int T = m->T, K = m->K;
    for (int t = 0; t < T; t++) {
        MC_TRY(dalloc(m, &m->d_logit[t], (size_t)m->n_qd * 4));
        MC_TRY(dalloc(m, &m->d_am[t], (size_t)m->n_qd * 4));
        MC_TRY(dalloc(m, &m->d_av[t], (size_t)m->n_qd * 4));
        MC_TRY(dalloc(m, &m->d_grad[t], (size_t)m->n_qd * 4));
    }
    MC_TRY(dalloc(m, &m->d_sh_logit, (size_t)K * 4));
    MC_TRY(dalloc(m, &m->d_sh_am, (size_t)K * 4));
    MC_TRY(dalloc(m, &m->d_sh_av, (size_t)K * 4));
    MC_TRY(dalloc(m, &m->d_sh2, (size_t)4 * kShAccStride * 4));     // alternate logit, m, v, value
    MC_TRY(dalloc(m, &m->d_gsh, (size_t)3 * kShAccSlot * 8));       // three steps of shared-table sums
    MC_HIP(hipMemset(m->d_gsh, 0, (size_t)3 * kShAccSlot * 8));
    MC_TRY(dalloc(m, &m->d_sh_grad, (size_t)K * 4));
    MC_TRY(dalloc(m, &m->d_sh_prior, (size_t)K * 8));
    MC_TRY(dalloc(m, &m->d_ll_part, (size_t)m->n_items * 8));
    MC_TRY(dalloc(m, &m->d_steps, 4));
    MC_HIP(hipMemset(m->d_steps, 0, 4));
    // Prior weights of the shared tables: 1 + the sessions that leave the
    // slot's temporary at (1,2) (the phantom count in sh_base).
    std::vector<double> base(2 * (size_t)K), prior((size_t)K);
    MC_HIP(hipMemcpy(base.data(), m->d_sh_base, base.size() * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < K; k++) {
        // PBM/UBM: the constant denominator of slot k is 2 + w + 2 ph with
        // w writers and ph = N - w sessions leaving it at (1,2), so
        // ph = den - 2 - N.
        double ph = 0.0;
        if (m->type == MC_PBM || m->type == MC_UBM) ph = base[2 * (size_t)k + 1] - 2.0 - (double)m->N;
        prior[(size_t)k] = 1.0 + ph;
    }
    MC_HIP(hipMemcpy(m->d_sh_prior, prior.data(), prior.size() * 8, hipMemcpyHostToDevice));
    MC_HIP(hipDeviceGetAttribute(&m->n_cu, hipDeviceAttributeMultiprocessorCount, m->device));
    m->grad_ready = true;
    return MC_OK;
}

GArgs make_gargs(mc_model* m, int batch) {
    GArgs G;
    std::memset(&G, 0, sizeof G);
    G.packed = m->d_packed;
    G.pidx16 = m->d_pidx16;
    G.pidx = m->d_pidx;
    G.mask = m->d_mask;
    G.len = m->varlen ? m->d_len : nullptr;
    G.N = m->N;
    G.n_w = m->n_w;
    G.n_pk = m->n_pk;
    G.items = m->d_items;
    G.lists = m->d_lists;
    for (int t = 0; t < 2; t++) {
        const int tt = t < m->T ? t : 0;
        G.val[t] = m->d_val[m->cur][tt];
        G.logit[t] = m->d_logit[tt];
        G.am[t] = m->d_am[tt];
        G.av[t] = m->d_av[tt];
        G.grad[t] = m->d_grad[tt];
        G.gacc[t] = m->d_gacc[tt];
    }
    G.counts = m->d_counts;
    G.counts16 = m->d_counts16;
    G.sh_val = m->d_sh_val[m->cur];
    G.sh_num = m->d_sh_num[m->cur];
    G.sh_den = m->d_sh_den[m->cur];
    G.sh_logit = m->d_sh_logit;
    G.sh_am = m->d_sh_am;
    G.sh_av = m->d_sh_av;
    G.wr_logit = m->d_sh_logit;
    G.wr_am = m->d_sh_am;
    G.wr_av = m->d_sh_av;
    G.wr_val = m->d_sh_val[m->cur];
    G.wr_main = 1;
    G.prev_batch = -1;
    G.sh_grad = m->d_sh_grad;
    G.sh_prior = m->d_sh_prior;
    G.sh_acc = m->d_gsh;
    G.ll_part = m->d_ll_part;
    G.shared_idx = m->d_shared_idx;
    G.steps = m->d_steps;
    G.K = m->K;
    G.kacc = m->KACC;
    G.max_width = m->max_width;
    G.batch = batch < 0 ? 0 : batch;
    G.n_batches = m->n_batches;
    G.use_prior = m->adam_prior;
    G.lr = m->adam_lr;
    G.b1 = m->adam_b1;
    G.b2 = m->adam_b2;
    G.eps = m->adam_eps;
    const int64_t nb = batch < 0 ? m->N : m->batch_sessions[(size_t)batch];
    G.sh_scale = nb > 0 ? (double)m->N / (double)nb : 1.0;
    return G;
}

template <int MODEL, bool WIN, bool ADAM, int NT>
void launch_items(mc_model* m, const GArgs& G, int64_t count, int64_t off, size_t lds) {
    if (m->varlen)
        hipLaunchKernelGGL((grad_item_kernel<MODEL, true, WIN, ADAM, NT>), dim3((unsigned)count), dim3(NT), lds,
                           m->stream, G, off);
    else
        hipLaunchKernelGGL((grad_item_kernel<MODEL, false, WIN, ADAM, NT>), dim3((unsigned)count), dim3(NT), lds,
                           m->stream, G, off);
}

template <int MODEL, bool ADAM>
void launch_grad_t(mc_model* m, const GArgs& G0, int batch) {
    const size_t b = batch < 0 ? (size_t)m->n_batches : (size_t)batch;
    const int64_t nw = m->win_cnt[b], ng = m->glob_cnt[b];
    GArgs Gb = G0;
    Gb.shared_off = m->shared_off[b];
    Gb.shared_cnt = m->shared_cnt[b];
    const GArgs& G = Gb;
    // Few items (a small minibatch): 1024 lanes per item to fill the chip.
    const bool wide = nw + ng < 4 * (int64_t)m->n_cu;
    if (nw > 0) {
        if (wide) launch_items<MODEL, true, ADAM, 1024>(m, G, nw, m->win_off[b], m->lds_bytes);
        else launch_items<MODEL, true, ADAM, 256>(m, G, nw, m->win_off[b], m->lds_bytes);
    }
    if (ng > 0) {
        if (wide) launch_items<MODEL, false, ADAM, 1024>(m, G, ng, m->glob_off[b], 0);
        else launch_items<MODEL, false, ADAM, 256>(m, G, ng, m->glob_off[b], 0);
    }
    const unsigned grid = (unsigned)(1 + (Gb.shared_cnt + 255) / 256);
    hipLaunchKernelGGL((grad_merge_kernel<MODEL, ADAM>), dim3(grid), dim3(256), 0, m->stream, Gb);
    if (ADAM && batch == m->n_batches - 1) hipLaunchKernelGGL(advance_epoch_kernel, dim3(1), dim3(64), 0, m->stream, G.steps);
}

// One Adam step of an epoch replayed from a graph: the step of batch b also
// finalises the step of batch b - 1 (its split entries and the shared
// tables), so an epoch is one launch per step plus a closing merge.  The
// shared tables' Adam state alternates between the model's arrays (version
// A) and d_sh2 (version B): step j >= 1 writes version j % 2 == 1 ? B : A,
// and the closing merge writes A.  Shared sums of step j go to d_gsh slot
// j % 3; step j clears slot (j + 1) % 3, last read by step j - 1.
template <int MODEL>
void launch_fused_step_t(mc_model* m, int b) {
    const int B = m->n_batches;
    float* A4[4] = {m->d_sh_logit, m->d_sh_am, m->d_sh_av, m->d_sh_val[m->cur]};
    float* B4[4] = {m->d_sh2, m->d_sh2 + kShAccStride, m->d_sh2 + 2 * kShAccStride, m->d_sh2 + 3 * kShAccStride};
    auto version = [&](int j) { return (j % 2 == 1) ? B4 : A4; };
    GArgs G = make_gargs(m, b);
    G.shared_off = m->shared_off[(size_t)b];
    G.shared_cnt = m->shared_cnt[(size_t)b];
    if (b < B) {
        float** rd = version(b > 0 ? b - 1 : 0);
        float** wr = version(b);
        G.sh_logit = rd[0];
        G.sh_am = rd[1];
        G.sh_av = rd[2];
        G.sh_val = rd[3];
        G.wr_logit = wr[0];
        G.wr_am = wr[1];
        G.wr_av = wr[2];
        G.wr_val = wr[3];
        G.wr_main = wr == A4;
        G.sh_acc = m->d_gsh + (size_t)(b % 3) * kShAccSlot;
        if (b > 0) {
            G.prev_batch = b - 1;
            G.prev_off = m->shared_off[(size_t)b - 1];
            G.prev_cnt = m->shared_cnt[(size_t)b - 1];
            const int64_t nb = m->batch_sessions[(size_t)b - 1];
            G.prev_scale = nb > 0 ? (double)m->N / (double)nb : 1.0;
            G.prev_acc = m->d_gsh + (size_t)((b - 1) % 3) * kShAccSlot;
            G.zero_acc = m->d_gsh + (size_t)((b + 1) % 3) * kShAccSlot;
        }
        const size_t bi = (size_t)b;
        const int64_t nw = m->win_cnt[bi], ng = m->glob_cnt[bi];
        const bool wide = nw + ng < 4 * (int64_t)m->n_cu;
        if (nw > 0) {
            if (wide) launch_items<MODEL, true, true, 1024>(m, G, nw, m->win_off[bi], m->lds_bytes);
            else launch_items<MODEL, true, true, 256>(m, G, nw, m->win_off[bi], m->lds_bytes);
        }
        if (ng > 0) {
            // A second launch of the same step: the finalisation is done and
            // the shared values are in the written version.
            if (nw > 0) {
                G.prev_batch = -1;
                G.sh_val = G.wr_val;
            }
            if (wide) launch_items<MODEL, false, true, 1024>(m, G, ng, m->glob_off[bi], 0);
            else launch_items<MODEL, false, true, 256>(m, G, ng, m->glob_off[bi], 0);
        }
        return;
    }
    // b == B: the closing merge of batch B - 1, into version A.
    float** rd = version(B - 1);
    G = make_gargs(m, B - 1);
    G.shared_off = m->shared_off[(size_t)B - 1];
    G.shared_cnt = m->shared_cnt[(size_t)B - 1];
    G.sh_logit = rd[0];
    G.sh_am = rd[1];
    G.sh_av = rd[2];
    G.sh_val = rd[3];
    G.sh_acc = m->d_gsh + (size_t)((B - 1) % 3) * kShAccSlot;
    G.zero_acc = m->d_gsh;
    const unsigned grid = (unsigned)(1 + (G.shared_cnt + 255) / 256);
    hipLaunchKernelGGL((grad_merge_kernel<MODEL, true>), dim3(grid), dim3(256), 0, m->stream, G);
    hipLaunchKernelGGL(advance_epoch_kernel, dim3(1), dim3(64), 0, m->stream, G.steps);
}

mc_status launch_fused_step(mc_model* m, int b) {
    switch (m->type) {
        case MC_PBM: launch_fused_step_t<kPBM>(m, b); break;
        case MC_UBM: launch_fused_step_t<kUBM>(m, b); break;
        case MC_DBN: launch_fused_step_t<kDBN>(m, b); break;
        default: launch_fused_step_t<kCCM>(m, b); break;
    }
    MC_HIP(hipGetLastError());
    return MC_OK;
}

mc_status launch_grad(mc_model* m, int batch, bool adam) {
    const GArgs G = make_gargs(m, batch);
    switch (m->type) {
        case MC_PBM: adam ? launch_grad_t<kPBM, true>(m, G, batch) : launch_grad_t<kPBM, false>(m, G, batch); break;
        case MC_UBM: adam ? launch_grad_t<kUBM, true>(m, G, batch) : launch_grad_t<kUBM, false>(m, G, batch); break;
        case MC_DBN: adam ? launch_grad_t<kDBN, true>(m, G, batch) : launch_grad_t<kDBN, false>(m, G, batch); break;
        default: adam ? launch_grad_t<kCCM, true>(m, G, batch) : launch_grad_t<kCCM, false>(m, G, batch); break;
    }
    MC_HIP(hipGetLastError());
    return MC_OK;
}

void mark_value_only(mc_model* m) {
    for (int t = 0; t < m->T; t++) m->den_state[t] = 3;
}

}  // namespace
}  // namespace mc

namespace mc {
namespace {

// Captures the graph of one epoch (B fused steps from batch 0 and the closing
// merge) if the current one does not start from the current buffers.
mc_status ensure_adam_graph(mc_model* m) {
    if (m->adam_graph && m->adam_graph_first == m->cur) return MC_OK;
    if (m->adam_graph) {
        (void)hipGraphExecDestroy(m->adam_graph);
        m->adam_graph = nullptr;
    }
    hipGraph_t g;
    MC_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    mc_status st = MC_OK;
    for (int b = 0; b <= m->n_batches && st == MC_OK; b++) st = launch_fused_step(m, b);
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    if (st != MC_OK) return st;
    MC_HIP(e);
    e = hipGraphInstantiate(&m->adam_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MC_HIP(e);
    m->adam_graph_first = m->cur;
    return MC_OK;
}

// Enqueues n_steps minibatch steps on the model's stream: whole epochs from
// batch 0 replay the epoch graph, other steps are launched one by one.
mc_status enqueue_adam_steps(mc_model* m, int n_steps, bool use_graph) {
    const int B = m->n_batches;
    int done = 0;
    while (use_graph && m->adam_next_batch == 0 && n_steps - done >= B) {
        MC_TRY(ensure_adam_graph(m));
        MC_HIP(hipGraphLaunch(m->adam_graph, m->stream));
        done += B;
    }
    for (; done < n_steps; done++) {
        MC_TRY(launch_grad(m, m->adam_next_batch, true));
        m->adam_next_batch = (m->adam_next_batch + 1) % B;
    }
    return MC_OK;
}

}  // namespace
}  // namespace mc

using namespace mc;

extern "C" {

mc_status mc_grad_init(mc_model* m) {
    MC_TRY(check_grad_support(m));
    MC_HIP(hipSetDevice(m->device));
    MC_TRY(ensure_grad_buffers(m));
    for (int t = 0; t < m->T; t++)
        hipLaunchKernelGGL(logit_init_kernel, dim3((unsigned)((m->n_qd + 255) / 256)), dim3(256), 0, m->stream,
                           m->d_val[m->cur][t], m->d_logit[t], m->d_am[t], m->d_av[t], m->n_qd);
    hipLaunchKernelGGL(logit_init_kernel, dim3(1), dim3(256), 0, m->stream, m->d_sh_val[m->cur], m->d_sh_logit,
                       m->d_sh_am, m->d_sh_av, (int64_t)m->K);
    hipLaunchKernelGGL(zero_steps_kernel, dim3(1), dim3(64), 0, m->stream, m->d_steps, 1);
    MC_HIP(hipGetLastError());
    MC_HIP(hipStreamSynchronize(m->stream));
    m->adam_next_batch = 0;
    return MC_OK;
}

mc_status mc_loglik_gradient(mc_model* m, int batch, double* loglik, float* const* grad) {
    MC_TRY(check_grad_support(m));
    MC_REQUIRE(batch >= -1 && batch < m->n_batches, "batch out of range");
    MC_HIP(hipSetDevice(m->device));
    MC_TRY(ensure_grad_buffers(m));
    for (int t = 0; t < m->T; t++) MC_TRY(fill(m, m->d_grad[t], 0.f, m->n_qd));
    MC_TRY(fill(m, (float*)m->d_ll_part, 0.f, 2 * (int64_t)m->n_items));
    MC_TRY(launch_grad(m, batch, false));
    MC_HIP(hipStreamSynchronize(m->stream));
    if (loglik) {
        std::vector<double> part((size_t)m->n_items);
        MC_HIP(hipMemcpy(part.data(), m->d_ll_part, part.size() * 8, hipMemcpyDeviceToHost));
        double s = 0.0;
        for (double v : part) s += v;
        *loglik = s;
    }
    if (grad) {
        for (int t = 0; t < m->T; t++)
            if (grad[t]) MC_HIP(hipMemcpy(grad[t], m->d_grad[t], (size_t)m->n_qd * 4, hipMemcpyDeviceToHost));
        if (grad[m->T]) MC_HIP(hipMemcpy(grad[m->T], m->d_sh_grad, (size_t)m->K * 4, hipMemcpyDeviceToHost));
    }
    return MC_OK;
}

mc_status mc_adam_configure(mc_model* m, float lr, float beta1, float beta2, float eps, int use_prior) {
    MC_REQUIRE(m, "null model");
    MC_REQUIRE(lr > 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f,
               "invalid Adam hyper-parameters");
    m->adam_lr = lr;
    m->adam_b1 = beta1;
    m->adam_b2 = beta2;
    m->adam_eps = eps;
    m->adam_prior = use_prior != 0;
    if (m->adam_graph) {
        (void)hipGraphExecDestroy(m->adam_graph);
        m->adam_graph = nullptr;
    }
    return MC_OK;
}

mc_status mc_adam_steps(mc_model* m, int n_steps, int use_graph) {
    MC_TRY(check_grad_support(m));
    MC_REQUIRE(m->grad_ready, "call mc_grad_init first");
    MC_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
    MC_HIP(hipSetDevice(m->device));
    MC_TRY(enqueue_adam_steps(m, n_steps, use_graph != 0));
    MC_HIP(hipStreamSynchronize(m->stream));
    mark_value_only(m);
    return MC_OK;
}

// Device time of n_steps steps: the epoch graph is captured before the first
// event, so only the replayed work is timed (as for mc_em_graph_timed).
mc_status mc_adam_steps_timed(mc_model* m, int n_steps, double* ms) {
    MC_REQUIRE(m && ms, "null argument");
    MC_TRY(check_grad_support(m));
    MC_REQUIRE(m->grad_ready, "call mc_grad_init first");
    MC_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
    MC_HIP(hipSetDevice(m->device));
    if (m->adam_next_batch == 0 && n_steps >= m->n_batches) MC_TRY(ensure_adam_graph(m));
    hipEvent_t a, b;
    MC_HIP(hipEventCreate(&a));
    MC_HIP(hipEventCreate(&b));
    MC_HIP(hipEventRecord(a, m->stream));
    mc_status st = enqueue_adam_steps(m, n_steps, true);
    MC_HIP(hipEventRecord(b, m->stream));
    MC_HIP(hipEventSynchronize(b));
    float t = 0.f;
    MC_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    if (st == MC_OK) mark_value_only(m);
    return st;
}

mc_status mc_get_logits(mc_model* m, int t, float* logits) {
    MC_REQUIRE(m && m->initialized && logits && t >= 0 && t <= m->T, "invalid argument");
    MC_REQUIRE(m->grad_ready, "call mc_grad_init first");
    MC_HIP(hipSetDevice(m->device));
    MC_HIP(hipStreamSynchronize(m->stream));
    const bool sh = t == m->T;
    MC_HIP(hipMemcpy(logits, sh ? m->d_sh_logit : m->d_logit[t], (size_t)(sh ? m->K : m->n_qd) * 4,
                     hipMemcpyDeviceToHost));
    return MC_OK;
}

mc_status mc_set_logits(mc_model* m, int t, const float* logits) {
    MC_REQUIRE(m && m->initialized && logits && t >= 0 && t <= m->T, "invalid argument");
    MC_REQUIRE(m->grad_ready, "call mc_grad_init first");
    MC_HIP(hipSetDevice(m->device));
    const bool sh = t == m->T;
    const int64_t n = sh ? m->K : m->n_qd;
    for (int64_t i = 0; i < n; i++)
        MC_REQUIRE(std::isfinite(logits[i]) && std::fabs(logits[i]) <= kLogitMax, "logit out of [-13.8, 13.8]");
    float* dl = sh ? m->d_sh_logit : m->d_logit[t];
    float* dv = sh ? m->d_sh_val[m->cur] : m->d_val[m->cur][t];
    MC_HIP(hipMemcpy(dl, logits, (size_t)n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(sigmoid_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, dl, dv, n);
    MC_HIP(hipGetLastError());
    if (sh) {
        hipLaunchKernelGGL(sigmoid_kernel, dim3(1), dim3(256), 0, m->stream, dl, m->d_sh_num[m->cur], n);
        MC_TRY(fill(m, m->d_sh_den[m->cur], 1.f, n));
    } else {
        m->den_state[t] = 3;
    }
    MC_HIP(hipStreamSynchronize(m->stream));
    return MC_OK;
}

}  // extern "C"
