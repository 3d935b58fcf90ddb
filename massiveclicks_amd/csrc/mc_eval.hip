// mc_eval.hip -- held-out evaluation on the device: per-session conditional
// and full click probabilities (pbm.cu:376-439, ubm.cu:373-468,
// dbn.cu:611-683, ccm.cu:604-671), the SERP click probabilities
// (get_serp_probability, pbm.cu:351, ubm.cu:345, dbn.cu:572, ccm.cu:566),
// and the log-likelihood and perplexity totals (evaluation.cpp:25-66) in
// FP64.
//
// Three session layouts feed the same per-session arithmetic:
//   * the caller's arrays, session-major ([n][10] int32 pair indices and
//     [n][10] click bytes), uploaded per call (mc_evaluate,
//     mc_serp_probability);
//   * a held-out set kept resident on the device (mc_set_test_set): sessions
//     reordered so that the sessions of one query are adjacent (their value
//     gathers then share cache lines), with the permutation back to the
//     caller's order for the per-session outputs, stored as one 16-byte
//     record per session when every session's seen pairs lie within 255
//     consecutive indices (pairs are numbered per query in contiguous
//     ranges), else rank-major ([10][n] int32 pair indices, 16-bit click
//     masks).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"
#include "mc_kernels.h"
#include "mc_model.h"

namespace mc {

struct EvalArgs {
    // Session-major layout.
    const int32_t* pidx;     // [n][10]
    const uint8_t* clicks;   // [n][10]
    // Resident layouts: packed records or rank-major.
    const uint4* packed;     // [n]: offsets, clicks and length, base (load_packed_test)
    const int32_t* rpidx;    // [10][n]
    const uint16_t* mask;    // [n]
    const int32_t* perm;     // [n]: caller's index of the session at position i, or null
    const uint8_t* len;      // [n] or null (10 ranks)
    int64_t n;
    int64_t n_qd;            // pair indices outside [0, n_qd) read as unseen pairs
    const float* val0;       // attractiveness values
    const float* val1;       // satisfaction values (DBN)
    const float* sh;         // shared values
    float* logp;             // [n][10] or null
    float* full;             // [n][10] or null
    float* llh;              // [n] or null
    float* prob;             // [n][10] or null: get_serp_probability
    double* partial;         // [gridDim.x][21]: llh, ppl[10], count[10]
};

// Value of pair p, or PARAM_DEF_NUM / PARAM_DEF_DENOM = 1/2 for a pair unseen
// in training (p = -1).  Indices the host has not checked (device-resident
// inputs) are bounded here: any p outside [0, n_qd) reads as unseen.
__device__ __forceinline__ float pv_or_default(const float* v, int p, int64_t n_qd) {
    return (uint64_t)(uint32_t)p < (uint64_t)n_qd ? v[p] : 0.5f;
}

// One session's pair indices and clicks, session-major: the ten indices as
// five 8-byte loads (rows are 40 B, so 8-byte aligned), the click bytes as
// five 2-byte loads.
__device__ __forceinline__ void load_session_major(const EvalArgs& E, int64_t i, int (&p)[kL], uint32_t& c) {
    const int2* p2 = reinterpret_cast<const int2*>(E.pidx + i * kL);
#pragma unroll
    for (int u = 0; u < kL / 2; u++) {
        const int2 v = p2[u];
        p[2 * u] = v.x;
        p[2 * u + 1] = v.y;
    }
    const uint16_t* c2 = reinterpret_cast<const uint16_t*>(E.clicks + i * kL);
    c = 0;
#pragma unroll
    for (int u = 0; u < kL / 2; u++) {
        const uint16_t v = c2[u];
        c |= (uint32_t)((v & 0xffu) == 1) << (2 * u);
        c |= (uint32_t)((v >> 8) == 1) << (2 * u + 1);
    }
}

// Rank-major: rank r of every session is one contiguous row, so a
// wavefront's 64 lanes read 256 contiguous bytes per rank.  Ranks past a
// session's length hold -1 and are read like the others, so no load waits
// for the length.
__device__ __forceinline__ void load_rank_major(const EvalArgs& E, int64_t i, int (&p)[kL], uint32_t& c) {
    c = E.mask[i];
#pragma unroll
    for (int r = 0; r < kL; r++) p[r] = E.rpidx[r * E.n + i];
}

// Packed test record: bytes 0-9 the offset of rank r's pair from the base
// (255: a pair unseen in training, or a rank past the length), bits 16-25
// of the third word the click bits and bits 26-29 the length, the fourth
// word the base (the session's smallest seen pair index).  One 16-byte load
// per session.
constexpr uint32_t kTestUnseen = 255u;
__device__ __forceinline__ void decode_packed_test(uint4 v, int& L, int (&p)[kL], uint32_t& c) {
    c = (v.z >> 16) & 0x3ffu;
    L = min(max((int)((v.z >> 26) & 0xfu), 1), kL);
    const uint32_t w[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int r = 0; r < kL; r++) {
        const uint32_t b = (w[r >> 2] >> (8 * (r & 3))) & 0xffu;
        p[r] = b == kTestUnseen ? -1 : (int)(v.w + b);
    }
}

// Session layouts of eval_kernel.
enum { kEvalSessionMajor = 0, kEvalRankMajor = 1, kEvalPacked = 2 };

// The per-session arithmetic, the reference's operation for operation in
// FP32: log conditional click probabilities lp (log2 instead of the natural
// log when !NATLOG, see log_click), full click probabilities
// fp and the SERP click probabilities sp of the first L ranks; click bits cm.
// The session's parameter values: attractiveness a and (DBN) satisfaction
// sv of its first L ranks.
template <int MODEL>
__device__ __forceinline__ void gather_values(const EvalArgs& E, int L, const int (&p)[kL], float (&a)[kL],
                                              float (&sv)[kL]) {
#pragma unroll
    for (int r = 0; r < kL; r++) {
        a[r] = r < L ? pv_or_default(E.val0, p[r], E.n_qd) : 0.5f;
        sv[r] = (MODEL == kDBN && r < L) ? pv_or_default(E.val1, p[r], E.n_qd) : 0.5f;
    }
}

// Log of a conditional click probability: the natural log (the reference's
// value, for per-session outputs) or, when only the sums are wanted, log2,
// which costs half the instructions; the caller scales a session's log2 sum
// by ln 2 once.
template <bool NATLOG>
__device__ __forceinline__ float log_click(float x) { return NATLOG ? logf(x) : log2f(x); }

template <int MODEL, bool NATLOG>
__device__ __forceinline__ void eval_session(const EvalArgs& E, int L, const float (&a)[kL], const float (&sv)[kL],
                                             uint32_t cm, float (&lp)[kL], float (&fp)[kL], float (&sp)[kL]) {
    bool c[kL];
#pragma unroll
    for (int r = 0; r < kL; r++) c[r] = (cm >> r) & 1u;
    if (MODEL == kPBM) {
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float ae = a[r] * E.sh[r];
                lp[r] = c[r] ? log_click<NATLOG>(ae) : log_click<NATLOG>(1 - ae);
                fp[r] = c[r] ? ae : 1 - ae;
                sp[r] = ae;
            }
        }
    } else if (MODEL == kUBM) {
        int prev[kL];
        int last = kL - 1;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            prev[r] = last;
            if (r < L && c[r]) last = r;
        }
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float ae = a[r] * E.sh[r * (r + 1) / 2 + prev[r]];
                lp[r] = c[r] ? log_click<NATLOG>(ae) : log_click<NATLOG>(1 - ae);
                sp[r] = ae;
            }
        }
        // P(C_r = 1) summed over the rank of the previous click (-1: none).
        // The reference forms each no-click product nb(rp, r) =
        // prod_{rb = rp+1}^{r-1} (1 - a_rb e[rb][rp]) afresh from 1.f in
        // increasing rb; nb[rp + 1] below keeps the running product of each
        // rp across r, which performs the same multiplications in the same
        // order (the same floats) in O(L^2) instead of O(L^3).
        float tf[kL];
        float nb[kL];
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                nb[r] = 1.f;   // rp = r - 1: empty product
                float cp = 0.f;
#pragma unroll
                for (int rp = -1; rp < r; rp++) {
                    const int crp = rp == -1 ? 9 : rp;
                    float temp = nb[rp + 1] * (a[r] * E.sh[r * (r + 1) / 2 + crp]);
                    cp += rp >= 0 ? tf[rp] * temp : temp;
                }
                tf[r] = cp;
                fp[r] = c[r] ? cp : 1 - cp;
#pragma unroll
                for (int rp = -1; rp < r; rp++) {
                    const int crp = rp == -1 ? 9 : rp;
                    nb[rp + 1] *= 1 - a[r] * E.sh[r * (r + 1) / 2 + crp];
                }
            } else {
                tf[r] = 0.f;
            }
        }
    } else if (MODEL == kDBN) {
        const float g = E.sh[0];
        float ex = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float cp;
                if (c[r]) { cp = a[r] * ex; ex = g * (1 - sv[r]); }
                else { cp = 1 - a[r] * ex; ex *= g * (1 - a[r]) / cp; }
                lp[r] = log_click<NATLOG>(cp);
                sp[r] = cp;
            }
        }
        ex = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float ae = a[r] * ex;
                fp[r] = c[r] ? ae : 1 - ae;
                ex *= g * (1 - a[r]) + g * a[r] * (1 - sv[r]);
            }
        }
    } else {
        const float t1 = E.sh[0], t2 = E.sh[1], t3 = E.sh[2];
        float ex = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float cp;
                if (c[r]) { cp = a[r] * ex; ex = t2 * (1 - a[r]) + t3 * a[r]; }
                else { cp = 1 - a[r] * ex; ex *= t1 * (1 - a[r]) / cp; }
                lp[r] = log_click<NATLOG>(cp);
                sp[r] = cp;
            }
        }
        ex = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                float ae = a[r] * ex;
                fp[r] = c[r] ? ae : 1 - ae;
                ex *= (1 - a[r]) * t1 + a[r] * ((1 - a[r]) * t2 + a[r] * t3);
            }
        }
    }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One session per lane, grid-stride: in every pass a wavefront takes 64
// consecutive sessions.  Per-block FP64 partials of the 21 sums; the per-rank
// session counts come from wavefront ballots (uniform values, no vector
// registers).  OUT: per-session outputs requested (separate instantiation,
// so that the common case keeps fewer values live).
template <int MODEL, int LAYOUT, bool OUT>
__global__ void __launch_bounds__(256) eval_kernel(EvalArgs E) {
    __shared__ double red[21][4];
    const int tid = threadIdx.x;
    // Per-lane FP32 partials of the lane's sessions (a lane takes every
    // (blocks x 256)-th session, about four of 2M test sessions), FP64 from
    // the wavefront sums on.
    float v[11];
#pragma unroll
    for (int k = 0; k < 11; k++) v[k] = 0.f;
    int cnt[kL];
#pragma unroll
    for (int r = 0; r < kL; r++) cnt[r] = 0;
    // The next session's length, indices and clicks are in flight while a
    // session is evaluated.
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
    // Packed records stay raw (4 registers) until their session is reached.
    int Ln = kL;
    int pn[kL];
    uint32_t cn = 0;
    uint4 rn = make_uint4(0u, 0u, 0u, 0u);
    auto load = [&](int64_t j) {
        if (LAYOUT == kEvalPacked) {
            rn = E.packed[j];
            return;
        }
        Ln = E.len ? min(max((int)E.len[j], 1), kL) : kL;
        if (LAYOUT == kEvalRankMajor) load_rank_major(E, j, pn, cn);
        else load_session_major(E, j, pn, cn);
    };
    auto take = [&](int& L, uint32_t& c, float (&a)[kL], float (&sv)[kL]) {
        int p[kL];
        L = Ln;
        c = cn;
        if (LAYOUT == kEvalPacked) {
            decode_packed_test(rn, L, p, c);
        } else {
#pragma unroll
            for (int r = 0; r < kL; r++) p[r] = pn[r];
        }
        gather_values<MODEL>(E, L, p, a, sv);
    };
    if (i < E.n) load(i);
    for (; i < E.n; i += stride) {
        int L;
        uint32_t c;
        float a[kL], sv[kL];
        take(L, c, a, sv);
        if (i + stride < E.n) load(i + stride);
        float lp[kL], fp[kL], sp[kL];
        eval_session<MODEL, OUT>(E, L, a, sv, c, lp, fp, sp);
        float acc = 0.f;
#pragma unroll
        for (int r = 0; r < kL; r++)
            if (r < L) acc = acc + lp[r];
        if (!OUT) acc *= 0.693147180559945309f;   // log2 sum to natural log
        const float s = acc / L;                 // evaluation.cpp:36
        v[0] += s;
        const int64_t o = (OUT && LAYOUT != kEvalSessionMajor && E.perm) ? (int64_t)E.perm[i] : i;
        if (OUT && E.llh) E.llh[o] = s;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            cnt[r] += (int)__popcll(__ballot(r < L));
            if (r < L) v[1 + r] += log2f(fp[r]);   // evaluation.cpp:62
            if (OUT && E.logp) E.logp[o * kL + r] = r < L ? lp[r] : 0.f;
            if (OUT && E.full) E.full[o * kL + r] = r < L ? fp[r] : 0.f;
            if (OUT && E.prob) E.prob[o * kL + r] = r < L ? sp[r] : 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < 11; k++) {
        double w = wave_sum_d((double)v[k]);
        if ((tid & 63) == 0) red[k][tid >> 6] = w;
    }
    if ((tid & 63) == 0)
#pragma unroll
        for (int r = 0; r < kL; r++) red[11 + r][tid >> 6] = (double)cnt[r];
    __syncthreads();
    if (tid < 21) {
        double w = 0.0;
        for (int j = 0; j < (int)(blockDim.x >> 6); j++) w += red[tid][j];
        E.partial[(int64_t)blockIdx.x * 21 + tid] = w;
    }
}

// Sums block partials: one block per output, each thread a fixed strided
// subset, then a fixed tree; the result does not depend on scheduling.
__global__ void __launch_bounds__(256) eval_reduce_kernel(const double* partial, int n_blocks, double* out) {
    __shared__ double red[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double w = 0.0;
    for (int b = tid; b < n_blocks; b += 256) w += partial[(int64_t)b * 21 + k];
    red[tid] = w;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    if (tid == 0) out[k] = red[0];
}

// Evaluation blocks per launch: at most this many (grid-stride beyond; each
// lane then has its next session's loads in flight; 4096 blocks measured
// slower).
constexpr int64_t kEvalMaxBlocks = 2048;

// Blocks for n sessions of a model: the UBM kernel measured fastest at 1536
// (46.1 us against 49.1 us with 2048 on the 1.94M-session set), the others
// at or near 2048 (DBN 45.4 us; 45.9 with 1792, 46.7 with 1536, 48.0 with
// 1024; CCM 41.3 us, 42.7 with 1536; PBM 30.6 us, 30.1 with 1792).
int64_t eval_blocks(int64_t n, int model_type) {
    const int64_t cap = model_type == MC_UBM ? 1536 : kEvalMaxBlocks;
    return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap));
}

// Scratch of mc_evaluate_device, the entry point for callers that cannot
// allocate device memory: the per-block FP64 partials of one evaluation
// (kEvalMaxBlocks x 21 doubles, 344 KB per device, held until the process
// ends).  Every block writes all 21 of its partials before the reduction
// reads them, so nothing has to be cleared between calls.  Host code never
// passes `scratch` itself to a kernel: its device address on each GPU comes
// from hipGetSymbolAddress (eval_scratch_setup).
constexpr size_t kScratchDoubles = (size_t)kEvalMaxBlocks * 21;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Largest n mc_evaluate_device takes (the kernel strides over any n with at
// most kEvalMaxBlocks blocks; the bound keeps session indices in 31 bits).
constexpr int64_t kEvalDeviceMaxSessions = INT32_MAX;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

mc_status eval_scratch_setup(int device) {
    MC_REQUIRE(device >= 0 && device < kMaxDevices, "device ordinal out of range of the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return MC_OK;
    MC_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    MC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (double*)ptr;
    return MC_OK;
}

double* eval_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

template <int LAYOUT, bool OUT>
void launch_eval_out(int type, const EvalArgs& E, int64_t nb, hipStream_t s) {
    switch (type) {
        case MC_PBM: hipLaunchKernelGGL((eval_kernel<kPBM, LAYOUT, OUT>), dim3((unsigned)nb), dim3(256), 0, s, E); break;
        case MC_UBM: hipLaunchKernelGGL((eval_kernel<kUBM, LAYOUT, OUT>), dim3((unsigned)nb), dim3(256), 0, s, E); break;
        case MC_DBN: hipLaunchKernelGGL((eval_kernel<kDBN, LAYOUT, OUT>), dim3((unsigned)nb), dim3(256), 0, s, E); break;
        default: hipLaunchKernelGGL((eval_kernel<kCCM, LAYOUT, OUT>), dim3((unsigned)nb), dim3(256), 0, s, E); break;
    }
}

template <int LAYOUT>
void launch_eval(int type, const EvalArgs& E, int64_t nb, hipStream_t s) {
    if (E.logp || E.full || E.llh || E.prob) launch_eval_out<LAYOUT, true>(type, E, nb, s);
    else launch_eval_out<LAYOUT, false>(type, E, nb, s);
}

EvalArgs model_eval_args(const mc_model* m) {
    EvalArgs E;
    std::memset(&E, 0, sizeof(E));
    E.val0 = m->d_val[m->cur][0];
    E.val1 = m->T > 1 ? m->d_val[m->cur][1] : m->d_val[m->cur][0];
    E.sh = m->d_sh_val[m->cur];
    E.n_qd = m->n_qd;
    return E;
}

// Checks the caller's session arrays (session-major, host memory).
mc_status check_sessions(const mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks,
                         const uint8_t* len) {
    for (int64_t i = 0; i < n; i++) {
        const int l = len ? len[i] : kMaxSerp;
        MC_REQUIRE(l >= 1 && l <= kMaxSerp, "len must be in [1, 10] (session " + std::to_string(i) + ")");
        for (int r = 0; r < l; r++) {
            const int32_t p = pidx[i * kMaxSerp + r];
            MC_REQUIRE(p >= -1 && p < m->n_qd, "test pidx out of [-1, n_qd) (session " + std::to_string(i) + ")");
            MC_REQUIRE(clicks[i * kMaxSerp + r] <= 1, "clicks must be 0 or 1");
        }
    }
    return MC_OK;
}

void unpack_totals(const double (&out)[21], double* llh_sum, double* ppl_log2_sum, int64_t* rank_count) {
    if (llh_sum) *llh_sum = out[0];
    for (int r = 0; r < kMaxSerp; r++) {
        if (ppl_log2_sum) ppl_log2_sum[r] = out[1 + r];
        if (rank_count) rank_count[r] = (int64_t)out[11 + r];
    }
}

// Session-major evaluation of caller arrays in host memory, with temporary
// device buffers for this call.
mc_status evaluate_host_arrays(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                               float* logp, float* full, float* llh, float* prob, double (&out)[21]) {
    MC_HIP(hipSetDevice(m->device));
    const int64_t nb = eval_blocks(n, m->type);
    std::vector<void*> tmp;
    auto talloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
        tmp.push_back(p);
        return p;
    };
    auto cleanup = [&]() { for (void* p : tmp) (void)hipFree(p); };
    int32_t* d_p = (int32_t*)talloc((size_t)n * kMaxSerp * 4);
    uint8_t* d_c = (uint8_t*)talloc((size_t)n * kMaxSerp);
    uint8_t* d_l = len ? (uint8_t*)talloc((size_t)n) : nullptr;
    float* d_lp = logp ? (float*)talloc((size_t)n * kMaxSerp * 4) : nullptr;
    float* d_fp = full ? (float*)talloc((size_t)n * kMaxSerp * 4) : nullptr;
    float* d_llh = llh ? (float*)talloc((size_t)n * 4) : nullptr;
    float* d_pr = prob ? (float*)talloc((size_t)n * kMaxSerp * 4) : nullptr;
    double* d_part = (double*)talloc((size_t)nb * 21 * 8);
    double* d_out = (double*)talloc(21 * 8);
    if (!d_p || !d_c || (len && !d_l) || (logp && !d_lp) || (full && !d_fp) || (llh && !d_llh) || (prob && !d_pr) ||
        !d_part || !d_out) {
        cleanup();
        set_error("device allocation failed in the evaluation");
        return MC_ERR_OOM;
    }
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess && x != hipSuccess) e = x; };
    chk(hipMemcpyAsync(d_p, pidx, (size_t)n * kMaxSerp * 4, hipMemcpyHostToDevice, m->stream));
    chk(hipMemcpyAsync(d_c, clicks, (size_t)n * kMaxSerp, hipMemcpyHostToDevice, m->stream));
    if (len) chk(hipMemcpyAsync(d_l, len, (size_t)n, hipMemcpyHostToDevice, m->stream));
    EvalArgs E = model_eval_args(m);
    E.pidx = d_p;
    E.clicks = d_c;
    E.len = d_l;
    E.n = n;
    E.logp = d_lp;
    E.full = d_fp;
    E.llh = d_llh;
    E.prob = d_pr;
    E.partial = d_part;
    launch_eval<kEvalSessionMajor>(m->type, E, nb, m->stream);
    chk(hipGetLastError());
    hipLaunchKernelGGL(eval_reduce_kernel, dim3(21), dim3(256), 0, m->stream, d_part, (int)nb, d_out);
    chk(hipGetLastError());
    chk(hipMemcpyAsync(out, d_out, 21 * 8, hipMemcpyDeviceToHost, m->stream));
    if (logp) chk(hipMemcpyAsync(logp, d_lp, (size_t)n * kMaxSerp * 4, hipMemcpyDeviceToHost, m->stream));
    if (full) chk(hipMemcpyAsync(full, d_fp, (size_t)n * kMaxSerp * 4, hipMemcpyDeviceToHost, m->stream));
    if (llh) chk(hipMemcpyAsync(llh, d_llh, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    if (prob) chk(hipMemcpyAsync(prob, d_pr, (size_t)n * kMaxSerp * 4, hipMemcpyDeviceToHost, m->stream));
    chk(hipStreamSynchronize(m->stream));
    cleanup();
    if (e != hipSuccess) {
        set_error(std::string("evaluation: ") + hipGetErrorString(e));
        return MC_ERR_HIP;
    }
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

mc_status mc_evaluate(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                      float* logp, float* full, float* llh, double* llh_sum, double* ppl_log2_sum,
                      int64_t* rank_count) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(n >= 1 && pidx && clicks, "invalid argument");
    MC_TRY(check_sessions(m, n, pidx, clicks, len));
    double out[21];
    MC_TRY(evaluate_host_arrays(m, n, pidx, clicks, len, logp, full, llh, nullptr, out));
    unpack_totals(out, llh_sum, ppl_log2_sum, rank_count);
    return MC_OK;
}

mc_status mc_serp_probability(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                              float* prob) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(n >= 1 && pidx && clicks && prob, "invalid argument");
    MC_TRY(check_sessions(m, n, pidx, clicks, len));
    double out[21];
    return evaluate_host_arrays(m, n, pidx, clicks, len, nullptr, nullptr, nullptr, prob, out);
}

mc_status mc_evaluate_device(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                             float* logp, float* full, float* llh, double* totals) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(pidx && clicks && totals, "null argument");
    MC_REQUIRE(n >= 1 && n <= kEvalDeviceMaxSessions,
               "n out of range [1, " + std::to_string(kEvalDeviceMaxSessions) + "] for mc_evaluate_device");
    const int64_t nb = eval_blocks(n, m->type);
    MC_REQUIRE((size_t)nb * 21 <= kScratchDoubles, "internal: evaluation partials exceed the scratch array");
    double* part = eval_scratch(m->device);
    MC_REQUIRE(part, "internal: evaluation scratch not set up on this device");
    MC_HIP(hipSetDevice(m->device));
    EvalArgs E = model_eval_args(m);
    E.pidx = pidx;
    E.clicks = clicks;
    E.len = len;
    E.n = n;
    E.logp = logp;
    E.full = full;
    E.llh = llh;
    E.partial = part;
    launch_eval<kEvalSessionMajor>(m->type, E, nb, m->stream);
    MC_HIP(hipGetLastError());
    hipLaunchKernelGGL(eval_reduce_kernel, dim3(21), dim3(256), 0, m->stream, (const double*)part, (int)nb, totals);
    MC_HIP(hipGetLastError());
    return MC_OK;
}

mc_status mc_set_test_set(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(n >= 1 && pidx && clicks, "invalid argument");
    MC_REQUIRE(n <= INT32_MAX, "too many test sessions");
    MC_TRY(check_sessions(m, n, pidx, clicks, len));
    MC_HIP(hipSetDevice(m->device));
    // Order: by the smallest valid pair index of each session.  Pair indices
    // are assigned per query in contiguous ranges, so this groups the
    // sessions of one query together; sessions without a valid index go
    // last.  Counting sort: O(n + n_qd), stable.
    const int64_t n_qd = m->n_qd;
    std::vector<int32_t> key((size_t)n);
    std::vector<int64_t> start((size_t)n_qd + 2, 0);
    for (int64_t i = 0; i < n; i++) {
        const int l = len ? len[i] : kMaxSerp;
        int32_t k = (int32_t)n_qd;
        for (int r = 0; r < l; r++) {
            const int32_t p = pidx[i * kMaxSerp + r];
            if (p >= 0 && p < k) k = p;
        }
        key[(size_t)i] = k;
        start[(size_t)k + 1]++;
    }
    for (int64_t k = 0; k <= n_qd; k++) start[(size_t)k + 1] += start[(size_t)k];
    std::vector<int32_t> perm((size_t)n);
    for (int64_t i = 0; i < n; i++) perm[(size_t)start[(size_t)key[(size_t)i]]++] = (int32_t)i;
    // Packed records when every session's seen pairs span at most 255
    // indices (the base is the smallest, key[i]); else rank-major.
    bool packable = true;
    for (int64_t i = 0; i < n && packable; i++) {
        const int l = len ? len[i] : kMaxSerp;
        for (int r = 0; r < l; r++) {
            const int32_t p = pidx[i * kMaxSerp + r];
            if (p >= 0 && p < n_qd && (int64_t)p - key[(size_t)i] >= (int64_t)kTestUnseen) packable = false;
        }
    }
    std::vector<uint32_t> h_pk(packable ? (size_t)n * 4 : 0);
    std::vector<int32_t> h_p(packable ? 0 : (size_t)n * kMaxSerp);
    std::vector<uint16_t> h_m(packable ? 0 : (size_t)n);
    std::vector<uint8_t> h_l(len && !packable ? (size_t)n : 0);
    for (int64_t j = 0; j < n; j++) {
        const int64_t i = perm[(size_t)j];
        const int l = len ? len[i] : kMaxSerp;
        uint16_t mk = 0;
        for (int r = 0; r < l; r++)
            if (clicks[i * kMaxSerp + r]) mk |= (uint16_t)(1u << r);
        if (packable) {
            const uint32_t base = key[(size_t)i] < n_qd ? (uint32_t)key[(size_t)i] : 0u;
            uint8_t b[12];
            for (int r = 0; r < kMaxSerp; r++) {
                const int32_t p = r < l ? pidx[i * kMaxSerp + r] : -1;
                b[r] = (p >= 0 && p < n_qd) ? (uint8_t)(p - (int32_t)base) : (uint8_t)kTestUnseen;
            }
            const uint32_t meta = (uint32_t)mk | ((uint32_t)l << 10);
            b[10] = (uint8_t)(meta & 0xffu);
            b[11] = (uint8_t)(meta >> 8);
            std::memcpy(&h_pk[(size_t)j * 4], b, 12);
            h_pk[(size_t)j * 4 + 3] = base;
        } else {
            for (int r = 0; r < kMaxSerp; r++) h_p[(size_t)r * n + j] = r < l ? pidx[i * kMaxSerp + r] : -1;
            h_m[(size_t)j] = mk;
            if (len) h_l[(size_t)j] = (uint8_t)l;
        }
    }
    // Buffers of an earlier test set are replaced.
    for (void* p : {(void*)m->d_test_packed, (void*)m->d_test_pidx, (void*)m->d_test_mask, (void*)m->d_test_len,
                    (void*)m->d_test_perm, (void*)m->d_test_part}) {
        if (!p) continue;
        auto it = std::find(m->allocations.begin(), m->allocations.end(), p);
        if (it != m->allocations.end()) m->allocations.erase(it);
        (void)hipFree(p);
    }
    m->mem_bytes -= m->test_bytes;
    m->test_bytes = 0;
    m->d_test_packed = nullptr;
    m->d_test_pidx = nullptr;
    m->d_test_mask = nullptr;
    m->d_test_len = nullptr;
    m->d_test_perm = nullptr;
    m->d_test_part = nullptr;
    m->n_test = 0;
    const size_t before = m->mem_bytes;
    const int64_t nb = eval_blocks(n, m->type);
    if (packable) {
        MC_TRY(dalloc(m, &m->d_test_packed, h_pk.size() * 4));
    } else {
        MC_TRY(dalloc(m, &m->d_test_pidx, h_p.size() * 4));
        MC_TRY(dalloc(m, &m->d_test_mask, h_m.size() * 2));
        if (len) MC_TRY(dalloc(m, &m->d_test_len, h_l.size()));
    }
    MC_TRY(dalloc(m, &m->d_test_perm, perm.size() * 4));
    MC_TRY(dalloc(m, &m->d_test_part, ((size_t)nb * 21 + 21) * 8));
    m->test_bytes = m->mem_bytes - before;
    if (packable) {
        MC_HIP(hipMemcpyAsync(m->d_test_packed, h_pk.data(), h_pk.size() * 4, hipMemcpyHostToDevice, m->stream));
    } else {
        MC_HIP(hipMemcpyAsync(m->d_test_pidx, h_p.data(), h_p.size() * 4, hipMemcpyHostToDevice, m->stream));
        MC_HIP(hipMemcpyAsync(m->d_test_mask, h_m.data(), h_m.size() * 2, hipMemcpyHostToDevice, m->stream));
        if (len) MC_HIP(hipMemcpyAsync(m->d_test_len, h_l.data(), h_l.size(), hipMemcpyHostToDevice, m->stream));
    }
    MC_HIP(hipMemcpyAsync(m->d_test_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, m->stream));
    MC_HIP(hipStreamSynchronize(m->stream));
    m->n_test = n;
    return MC_OK;
}

mc_status mc_evaluate_test_set(mc_model* m, float* logp, float* full, float* llh, double* llh_sum,
                               double* ppl_log2_sum, int64_t* rank_count, double* kernel_ms) {
    MC_REQUIRE(m && m->initialized, "model not initialised");
    MC_REQUIRE(m->n_test > 0, "no test set: call mc_set_test_set first");
    MC_HIP(hipSetDevice(m->device));
    const int64_t n = m->n_test;
    const int64_t nb = eval_blocks(n, m->type);
    std::vector<void*> tmp;
    auto cleanup = [&]() { for (void* p : tmp) (void)hipFree(p); };
    auto talloc = [&](size_t bytes) -> float* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
        tmp.push_back(p);
        return (float*)p;
    };
    float* d_lp = logp ? talloc((size_t)n * kMaxSerp * 4) : nullptr;
    float* d_fp = full ? talloc((size_t)n * kMaxSerp * 4) : nullptr;
    float* d_llh = llh ? talloc((size_t)n * 4) : nullptr;
    if ((logp && !d_lp) || (full && !d_fp) || (llh && !d_llh)) {
        cleanup();
        set_error("device allocation failed in mc_evaluate_test_set");
        return MC_ERR_OOM;
    }
    EvalArgs E = model_eval_args(m);
    E.packed = reinterpret_cast<const uint4*>(m->d_test_packed);
    E.rpidx = m->d_test_pidx;
    E.mask = m->d_test_mask;
    E.perm = m->d_test_perm;
    E.len = m->d_test_len;
    E.n = n;
    E.logp = d_lp;
    E.full = d_fp;
    E.llh = d_llh;
    E.partial = m->d_test_part;
    double* d_out = m->d_test_part + nb * 21;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess && x != hipSuccess) e = x; };
    if (kernel_ms) {
        chk(hipEventCreate(&ev[0]));
        chk(hipEventCreate(&ev[1]));
        chk(hipEventRecord(ev[0], m->stream));
    }
    if (m->d_test_packed) launch_eval<kEvalPacked>(m->type, E, nb, m->stream);
    else launch_eval<kEvalRankMajor>(m->type, E, nb, m->stream);
    chk(hipGetLastError());
    if (kernel_ms) chk(hipEventRecord(ev[1], m->stream));
    hipLaunchKernelGGL(eval_reduce_kernel, dim3(21), dim3(256), 0, m->stream, (const double*)m->d_test_part, (int)nb,
                       d_out);
    chk(hipGetLastError());
    double out[21];
    chk(hipMemcpyAsync(out, d_out, 21 * 8, hipMemcpyDeviceToHost, m->stream));
    if (logp) chk(hipMemcpyAsync(logp, d_lp, (size_t)n * kMaxSerp * 4, hipMemcpyDeviceToHost, m->stream));
    if (full) chk(hipMemcpyAsync(full, d_fp, (size_t)n * kMaxSerp * 4, hipMemcpyDeviceToHost, m->stream));
    if (llh) chk(hipMemcpyAsync(llh, d_llh, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    chk(hipStreamSynchronize(m->stream));
    if (kernel_ms && e == hipSuccess) {
        float ms = 0.f;
        chk(hipEventElapsedTime(&ms, ev[0], ev[1]));
        *kernel_ms = ms;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    cleanup();
    if (e != hipSuccess) {
        set_error(std::string("mc_evaluate_test_set: ") + hipGetErrorString(e));
        return MC_ERR_HIP;
    }
    unpack_totals(out, llh_sum, ppl_log2_sum, rank_count);
    return MC_OK;
}

}  // extern "C"
