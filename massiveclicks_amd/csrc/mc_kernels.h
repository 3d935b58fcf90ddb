// mc_kernels.h -- CDNA4 device code of the EM engine (PBM, UBM, CCM, DBN).
//
// One workgroup processes one work item: a contiguous run of training
// sessions (sorted by query, so the query-document entries it touches form a
// contiguous window).  Each lane runs one session at a time; the expected
// counts it produces are added into an LDS window of 64-bit fixed-point
// accumulators (2^-32 resolution), which makes the per-entry sums independent
// of the order in which lanes add them.  At the end of the item the window is
// flushed: entries owned by this item are finalised in place (theta = prior +
// sum), entries shared with other items (query groups split over several
// items) are merged with one 64-bit integer atomic per entry and finalised by
// the merge kernel.  The shared (rank-indexed) tables are reduced per work
// item and merged with one 64-bit integer atomic per accumulator.
//
// Expected counts that do not depend on the parameters (clicked ranks, the
// (1,2) contributions of temporary slots the reference never writes, the
// exact (1,1) gamma/tau1 terms of ranks before the last click) are folded
// into per-entry constants at init, so a lane only emits the counts that
// change between iterations.  The derivations are in the comments of each
// model's session function.
#ifndef MC_KERNELS_H
#define MC_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mc {

constexpr int kL = 10;                          // MAX_SERP, utils/definitions.h:50
constexpr float kClampMax = 1.f - 0.000001f;    // param.cu:79
constexpr float kFixScale = 4294967296.0f;      // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;

enum { kPBM = 0, kCCM = 1, kDBN = 2, kUBM = 3 };

// Work item flags.  kItemPacked: a window item whose sessions are stored as
// 16-byte records (every session's indices lie within 256 consecutive
// window entries).
enum { kItemInterior = 1, kItemDirect = 2, kItemPacked = 4 };

struct Item {
    int32_t s0, s1;     // session range [s0, s1) in device order
    int32_t base;       // first qd index of the accumulator window
    int32_t width;      // window width; 0 = no window (global atomics)
    int32_t flags;      // kItemInterior: every window entry is owned by this item;
                        // kItemDirect: every session owns its entries alone;
                        // kItemPacked: sessions stored as 16-byte records
    int32_t pad;        // minibatch of the gradient path
    int32_t e1;         // end of the sessions the EM kernels visit, s0 <= e1 <= s1: the
                        // sessions [e1, s1) contribute only constant counts (DBN and
                        // CCM sessions without a click) and are read by the gradient
                        // path alone
    int32_t split;      // slices of one query group: the group's index in
                        // KArgs::split_groups, else -1
};

struct TableArgs {
    const float* val_in;      // value() of theta_t, [n_qd]
    float* val_out;           // value() of theta_{t+1}
    float* num_out;           // numerator of theta_{t+1}
    unsigned long long* gacc; // fixed-point accumulators of shared entries, [n_qd]
    float* gaccf;             // float accumulators (atomic mode), [n_qd]
};

struct KArgs {
    // Sessions in device order, three formats:
    //   [0, n_pk)   packed window items: one 16-byte record per session
    //               (PackedSession below), indices relative to the window;
    //   [n_pk, n_w) other window items: 16-bit window-relative indices,
    //               rank-major;
    //   [n_w, N)    the rest: absolute 32-bit indices, rank-major.
    const uint32_t* packed;   // [n_pk][4]
    const uint16_t* pidx16;   // [kL][n_w - n_pk]
    const int32_t* pidx;      // [kL][N - n_w] ((N - n_w) * 40 bytes < 2^32)
    const uint16_t* mask;     // [N] click bits (bit r = rank r)
    const uint8_t* len;       // [N] ranks shown, or nullptr
    int64_t N;
    int64_t n_w;
    int64_t n_pk;
    const Item* items;
    int32_t n_items;
    int32_t mode;             // 0 segmented, 1 float atomics
    TableArgs tab[2];
    const int2* counts;       // per qd entry: (occurrences, clicked occurrences) in training
    const uint16_t* counts16; // the same packed as o | c << 8 when o < 255, else kCountsEscape
    const float* sh_val_in;   // shared table values of theta_t
    unsigned long long* sh_acc;   // [kacc] fixed-point accumulators of the shared tables
    int32_t max_width;            // largest item window (LDS sizing)
    // merge kernel
    const int32_t* shared_idx;   // entries merged across items
    int64_t n_shared;
    int64_t n_qd;
    int32_t n_tables;
    int32_t write_num;
    // shared finalisation
    int32_t K;                // shared table size
    int32_t kacc;             // shared accumulators
    const double* sh_base;    // [2*K]: base numerator, base denominator
    const int32_t* sh_acc_map;// [2*K]: accumulator index of num/den, -1 = none
    float* sh_val_out;
    float* sh_num_out;
    float* sh_den_out;
    // M-step inside the persistent kernel (fuse_merge): the last slice of a
    // split query group to finish finalises the group's entries, the last
    // workgroup the shared tables; counters return to 0 for the next launch.
    int32_t fuse_merge;
    const int4* split_groups;     // per split group: first entry, end entry, slices
    unsigned* split_done;         // per split group: slices finished
    unsigned* grid_done;          // workgroups finished
};

// Counts of an entry from its 2-byte packed form, or from the full array
// when they do not fit (o >= 255; clicked occurrences never exceed o).
constexpr uint32_t kCountsEscape = 0xFFFFu;
__device__ __forceinline__ int2 unpack_counts(uint32_t pk, const int2* counts, int p) {
    return pk != kCountsEscape ? make_int2((int)(pk & 0xFFu), (int)(pk >> 8)) : counts[p];
}
// The same with the full counts `full` already loaded (the persistent
// kernel's flush loads them for every entry slot, without a branch).
__device__ __forceinline__ int2 unpack_counts(uint32_t pk, int2 full) {
    return pk != kCountsEscape ? make_int2((int)(pk & 0xFFu), (int)(pk >> 8)) : full;
}

// Parameter-independent part of the M-step of per-qd table t for an entry
// with `o` occurrences of which `c` were clicked: (prior + constant expected
// counts, denominator).  Prior (1, 2) (utils/definitions.h:42-43); every
// occurrence adds 1 to the denominator, CCM clicks add 2 (ccm.cu:311-314);
// clicked occurrences add exactly 1 to the attractiveness numerator; for the
// DBN satisfaction, clicked ranks add 1 to the denominator and unclicked
// ranks keep the (1,2) of an unwritten temporary slot (dbn.cu:350-378).
template <int MODEL>
__device__ __forceinline__ float2 base_den(int t, int2 oc) {
    const float o = (float)oc.x, c = (float)oc.y;
    if (MODEL == kCCM) return make_float2(1.f + c, 2.f + o + c);
    if (MODEL == kDBN && t == 1) return make_float2(1.f + (o - c), 2.f + 2.f * o - c);
    return make_float2(1.f + c, 2.f + o);
}

__device__ __forceinline__ float clamp_value(float num, float den) {   // param.cu:75-87
    float v = num / den;
    return v < kClampMax ? v : kClampMax;
}

// Non-negative float -> 32.32 fixed point, truncating below 2^-32.  Every
// expected count is a probability ratio >= 0; the integer part and the
// fraction are converted separately with the native u32 conversion.
__device__ __forceinline__ long long to_fix(float x) {
    const uint32_t hi = (uint32_t)x;
    const float f = x - (float)hi;
    const uint32_t lo = (uint32_t)(f * kFixScale);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// to_fix for a posterior x in [0, 1]: every float below 1 is at most
// 1 - 2^-24, so x * 2^32 <= 2^32 - 256 converts exactly and the result equals
// to_fix(x).  v_cvt_u32_f32 saturates inputs above the u32 range to
// 2^32 - 1 (and NaN to 0), so x >= 1, which the hardware reciprocal can give
// for a posterior of exactly 1, becomes 1 - 2^-32 with no separate clamp;
// written as asm because a float-to-unsigned conversion out of range is
// undefined in C++.  Two instructions instead of five.
__device__ __forceinline__ unsigned long long to_fix_unit(float x) {
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x * kFixScale));
    return (unsigned long long)r;
}

// a / b through the hardware reciprocal (1 ulp); used in the DBN and CCM
// chains, where the reference's own arithmetic (log/exp of products) differs
// from these restructured products by more than this anyway.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Highest set bit of m, or 10 when m == 0 (search.cu:160-173).
__device__ __forceinline__ int last_click(uint32_t m) {
    return m ? 31 - __builtin_clz(m) : kL;
}

// LDS accumulator window of a work item: `wstride` slots per per-qd table,
// table t's slots contiguous (a float2-wide entry stride would halve the
// banks that one table's 8-byte atomics spread over).  Entry p of a window of
// W entries owns 2^rsh consecutive slots (copies), 2^rsh <= wstride / W:
// lane l adds into copy l mod 2^rsh, which spreads the atomics of lanes that
// hit one entry (the narrow windows of the heaviest queries) over several
// addresses, and the flush adds the copies.
template <int T>
struct AccWindow {
    unsigned long long* acc;
    int wstride;
    int rsh;
    __device__ __forceinline__ int slot(int t, int q) const { return t * wstride + q; }   // q = (p << rsh) + copy
    __device__ __forceinline__ unsigned long long sum(int t, int p) const {
        if (rsh == 0) return acc[slot(t, p)];
        unsigned long long v = 0;
        for (int k = 0; k < (1 << rsh); k++) v += acc[slot(t, (p << rsh) + k)];
        return v;
    }
};

// Copies per entry of a window of W entries in wstride slots: a power of
// two, at most 64.
__device__ __forceinline__ int acc_copies_log2(int W, int wstride) {
    if (W <= 0) return 0;
    const int q = wstride / W;
    return q <= 1 ? 0 : min(6, 31 - __builtin_clz((unsigned)q));
}

// --------------------------------------------------------------------------
// Emitters: where a per-qd expected count goes.
// --------------------------------------------------------------------------
// add(t, r, p, x): expected count x, a posterior in [0, 1], of per-qd table
// t for the entry p shown at rank r of the current session; add_wide takes
// any non-negative x (the CCM's last-click attractiveness term can exceed 1;
// the last-click terms, once per session, always take it).
// UBM examination counts of a workgroup in LDS: kUbmSlots fixed-point
// accumulators, each in kUbmCopies copies; lane l adds into copy
// l mod kUbmCopies, so the lanes of one instruction never share an address
// with more than one other lane.  Laid out [slot][copy].
constexpr int kUbmSlots = 55;
constexpr int kUbmCopies = 32;
constexpr size_t kUbmSlotBytes = (size_t)kUbmSlots * kUbmCopies * 8;

template <int T>
struct EmitLds {            // fixed-point LDS window, relative indices
    char* base[T];             // this lane's copy of entry 0 of each table: &acc[slot(t, copy)]
    int sh;                    // rsh + 3: entry p's copies start p << sh bytes further
    unsigned long long* slots; // UBM: this lane's copy of the examination slots (stride kUbmCopies)
    static constexpr bool kSlotsInLds = true;
    __device__ __forceinline__ EmitLds(const AccWindow<T>& w, int copy, unsigned long long* lane_slots)
        : sh(w.rsh + 3), slots(lane_slots) {
#pragma unroll
        for (int t = 0; t < T; t++) base[t] = (char*)&w.acc[w.slot(t, copy)];
    }
    // One shift-add per emission: base[t] + (p << sh) is &acc[slot(t, (p << rsh) + copy)].
    __device__ __forceinline__ unsigned long long* at(int t, int p) const {
        return (unsigned long long*)(base[t] + (p << sh));
    }
    __device__ __forceinline__ void add(int t, int, int p, float x) { atomicAdd(at(t, p), to_fix_unit(x)); }
    __device__ __forceinline__ void add_wide(int t, int, int p, float x) {
        atomicAdd(at(t, p), (unsigned long long)to_fix(x));
    }
    __device__ __forceinline__ void add_slot(int k, float x) { atomicAdd(&slots[k * kUbmCopies], to_fix_unit(x)); }
};
struct EmitGlobalFix {      // fixed-point global (item without a window)
    static constexpr bool kSlotsInLds = false;
    unsigned long long* g0;
    unsigned long long* g1;
    __device__ __forceinline__ void add(int t, int, int p, float x) {
        atomicAdd(t == 0 ? &g0[p] : &g1[p], to_fix_unit(x));
    }
    __device__ __forceinline__ void add_wide(int t, int, int p, float x) {
        atomicAdd(t == 0 ? &g0[p] : &g1[p], (unsigned long long)to_fix(x));
    }
};
template <int T>
struct EmitLdsFloat {       // float LDS window (float-atomic mode), relative indices
    static constexpr bool kSlotsInLds = false;
    float* acc;                // [width][T]
    __device__ __forceinline__ void add(int t, int, int p, float x) { atomicAdd(&acc[p * T + t], x); }
    __device__ __forceinline__ void add_wide(int t, int r, int p, float x) { add(t, r, p, x); }
};
struct EmitGlobalFloat {    // float atomics (reference-like mode)
    static constexpr bool kSlotsInLds = false;
    float* g0;
    float* g1;
    __device__ __forceinline__ void add(int t, int, int p, float x) {
        unsafeAtomicAdd(t == 0 ? &g0[p] : &g1[p], x);
    }
    __device__ __forceinline__ void add_wide(int t, int r, int p, float x) { add(t, r, p, x); }
};
// Sessions whose entries no other session touches (a query group of one
// session): the counts stay in registers, indexed by rank, and the lane
// writes the new parameter values itself.
struct EmitRegs {
    static constexpr bool kSlotsInLds = false;
    float x[2][kL];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int r = 0; r < kL; r++) x[t][r] = 0.f;
    }
    __device__ __forceinline__ void add(int t, int r, int, float v) { x[t][r] += v; }
    __device__ __forceinline__ void add_wide(int t, int r, int p, float v) { add(t, r, p, v); }
};

// --------------------------------------------------------------------------
// Value sources: where theta_t is read from.
// --------------------------------------------------------------------------
// Window items address LDS with the query-document index relative to the
// window base (the device copy of the indices is stored that way), with the
// T per-qd tables interleaved per entry.
template <int T>
struct SrcLds {             // the item's window of values, staged in LDS
    const float* v;            // [width][T]
    __device__ __forceinline__ float get(int t, int p) const { return v[p * T + t]; }
    __device__ __forceinline__ float2 pair(int p) const { return reinterpret_cast<const float2*>(v)[p]; }
};
struct SrcGlobal {          // straight from HBM, absolute indices
    const float* v0;
    const float* v1;
    __device__ __forceinline__ float get(int t, int p) const { return t == 0 ? v0[p] : v1[p]; }
    __device__ __forceinline__ float2 pair(int p) const { return make_float2(v0[p], v1[p]); }
};

// --------------------------------------------------------------------------
// Per-model session kernels.  Each takes the session's query-document
// indices pv (ranks < L valid), the rank count L, the click bits mm
// (restricted to ranks < L) and emits the counts that vary between
// iterations.
// --------------------------------------------------------------------------

// PBM E-step, pbm.cu:191-231.  Clicked ranks contribute (1,1) to both tables
// (constant); unclicked ranks contribute ((a-ae)/(1-ae), 1) and ((e-ae)/(1-ae), 1),
// both through one hardware reciprocal of 1 - ae (1 ulp).
template <class Src, class Emit>
__device__ __forceinline__ void pbm_session(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                            const float* sh, float (&acc)[55]) {
    // ae = a e enters 1 - ae, a - ae and e - ae as multiply-adds.
#pragma clang fp contract(fast)
#pragma unroll
    for (int r = 0; r < kL; r++) {
        if (r < L && !((mm >> r) & 1u)) {
            const int p = pv[r];
            float atr = src.get(0, p);
            float ex = sh[r];
            float ae = atr * ex;
            const float inv = __builtin_amdgcn_rcpf(1 - ae);
            float na = (atr - ae) * inv;
            float ne = (ex - ae) * inv;
            em.add(0, r, p, na);
            acc[r] += ne;
        }
    }
}

// Examination slot of each rank for the UBM: r(r+1)/2 + prev[r], prev = 9
// when no earlier click (ubm.cu:201, search.cu:184-198).
__device__ __forceinline__ void ubm_slots(uint32_t mm, int (&slot)[kL]) {
#pragma unroll
    for (int r = 0; r < kL; r++) {
        uint32_t before = mm & ((1u << r) - 1u);
        int prev = before ? 31 - __builtin_clz(before) : kL - 1;
        slot[r] = r * (r + 1) / 2 + prev;
    }
}

// UBM E-step, ubm.cu:185-223.  The "no earlier click" slot of a rank can
// coincide with a later rank's slot in the same session (e.g. no click
// before rank 2 -> slot 12; click at rank 2 -> rank 4 uses slot 10 + 2 = 12);
// the reference writes tmp[slot][session] in rank order, so only the last
// rank writing a slot contributes to it.  An unclicked rank r is always that
// last writer: a later rank q shares its slot r(r+1)/2 + 9 only when q's
// previous click c satisfies q(q+1)/2 + c = r(r+1)/2 + 9, which forces
// c <= r, while "no click before r" and r unclicked leave no click at or
// before r.  Only clicked ranks are ever overwritten, and their constant
// (1,1) terms are counted on the host.
template <class Src, class Emit>
__device__ __forceinline__ void ubm_session(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                            const float* sh, float (&acc)[55]) {
    // ae = a e enters 1 - ae, a - ae and e - ae as multiply-adds.
#pragma clang fp contract(fast)
    if (mm == 0) {
        // No click: every rank's slot is its row's "no previous click" one,
        // known at compile time, so the select over the row is not needed.
        // Sessions are sorted by last click inside an item, so click-free
        // sessions fill whole wavefronts.
#pragma unroll
        for (int r = 0; r < kL; r++) {
            if (r < L) {
                const int row = r * (r + 1) / 2;
                const int p = pv[r];
                const float atr = src.get(0, p);
                const float ex = sh[row + kL - 1];
                const float ae = atr * ex;
                const float inv = __builtin_amdgcn_rcpf(1 - ae);
                em.add(0, r, p, (atr - ae) * inv);
                acc[row + kL - 1] += (ex - ae) * inv;
            }
        }
        return;
    }
    // The previous click of rank r follows the ranks in order (kL - 1 until
    // the first click), so each rank's slot costs one select.
    int prev = kL - 1;
#pragma unroll
    for (int r = 0; r < kL; r++) {
        const bool clicked = (mm >> r) & 1u;
        const int row = r * (r + 1) / 2;
        const int slot = row + prev;
        if (r < L && !clicked) {
            const int p = pv[r];
            float atr = src.get(0, p);
            float ex = sh[slot];
            float ae = atr * ex;
            const float inv = __builtin_amdgcn_rcpf(1 - ae);
            float na = (atr - ae) * inv;
            float ne = (ex - ae) * inv;
            em.add(0, r, p, na);
            if constexpr (Emit::kSlotsInLds) {
                // Before the first click the slot is the row's "no earlier
                // click" one, known at compile time (a register, as for
                // click-free sessions); after it, one LDS add into the slot.
                if (prev == kL - 1) acc[row + kL - 1] += ne;
                else em.add_slot(slot, ne);
            } else {
#pragma unroll
                for (int j = 0; j < r; j++) acc[row + j] += (slot == row + j) ? ne : 0.f;
                acc[row + kL - 1] += (slot == row + kL - 1) ? ne : 0.f;
            }
        }
        prev = clicked ? r : prev;
    }
}

// Select v[i] for a runtime i in [0, 10) without dynamic register indexing.
__device__ __forceinline__ float pick10(const float (&v)[kL], int i) {
    float out = v[0];
#pragma unroll
    for (int k = 1; k < kL; k++) out = (i == k) ? v[k] : out;
    return out;
}

// DBN E-step, dbn.cu:207-467, for a session with at least one click.
//
// * compute_exm_car (dbn.cu:266-296): exam[r+1] = exam[r](t_r + g a_r (1-s_r)),
//   t_r = g(1-a_r); car[r] = a_r + t_r car[r+1], car[L] = 0.
// * attractiveness (dbn.cu:312-336): clicked ranks give (1,1) (constant),
//   unclicked ranks before the last click give (0,1) (constant), unclicked
//   ranks after it give a(1-exam)/(1-exam car).
// * satisfaction (dbn.cu:350-378): only the last clicked rank varies:
//   s/(1-(1-s) g car[lcr+1]); other clicked ranks give (0,1), unclicked
//   ranks keep the (1,2) of the never-written slot (constants).
// * gamma (dbn.cu:436-467 with the factor of factor.cu:206-274): for r < lcr
//   every factor but phi5 is masked to 0, so phi5/sum = (phi4+phi5)/sum = 1
//   exactly (constant (1,1)).  For r >= lcr the z=1 factors need the product
//   of the tail click probabilities of start rank r+1, and that chain
//   (dbn.cu:390-423) uses the parameters of ranks 0.. with the clicks of
//   ranks r+1.. (all unclicked): its product telescopes to
//   V_{L-1-r}, V_{k+1} = V_k - a_k G_k, G_{k+1} = G_k t_k, V_0 = G_0 = 1.
//   Factors common to all non-zero phi (a for the clicked rank, (1-a) and
//   the examination probability otherwise) cancel in the ratios.  The
//   conditional examination after the last click follows
//   E_{lcr+1} = g(1-s_lcr), E_{r+1} = E_r t_r / (1 - a_r E_r).
template <bool VARLEN, class Src, class Emit>
__device__ __forceinline__ void dbn_session(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                            float g, float& gn, float& gd) {
    // The chains are restructured algebra rather than the reference's
    // operation order, so multiply-adds may fuse here.
#pragma clang fp contract(fast)
    const int lcr = last_click(mm);
    // Pass 1, ranks 0..L-1: exam[r] (compute_exm_car) and the no-click tail
    // products V_k; only exam[r] for r > lcr and V_k for k < L - lcr are used.
    // The examination factor t + g a (1 - s) equals g (1 - a s): with
    // ga = g a, one multiply-add gives it and one subtraction gives t.
    float exam[kL], V[kL];
    {
        float e = 1.f, v = 1.f, G = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            exam[r] = e;
            V[r] = v;
            if (!VARLEN || r < L) {
                const float2 as = src.pair(pv[r]);
                const float ga = g * as.x;
                e = e * (g - ga * as.y);
                v = v - as.x * G;
                G = G * (g - ga);
            }
        }
    }
    // Pass 2, ranks L-1 down to lcr+1 (all unclicked): car[r] and the
    // attractiveness posterior a(1-exam)/(1-exam car).  The parameters are
    // read again from LDS instead of being kept live in registers across the
    // passes (the compiler barrier stops the loads from being merged).
    __asm__ volatile("" ::: "memory");
    float car = 0.f;
#pragma unroll
    for (int r = kL - 1; r >= 0; r--) {
        if ((!VARLEN || r < L) && r > lcr) {
            const float a = src.get(0, pv[r]);
            car = a + (g - g * a) * car;
            em.add(0, r, pv[r], fdiv(a - exam[r] * a, 1 - exam[r] * car));
        }
    }
    // Pass 3, ranks lcr..L-1: satisfaction at the last click, gamma factors.
    // After the last click the conditional examination probability E
    // follows E_{r+1} = E_r g(1-a_r) / d2, d2 = 1 - a_r E_r, and the factor
    // ratios of an unclicked rank are phi5/sum = g T E / d1 and
    // (phi4+phi5)/sum = E (1-g+gT) / d1, d1 = 1 - g E (1-T).  Both divisions
    // share one reciprocal q = 1/(d1 d2): 1/d1 = d2 q, 1/d2 = d1 q.  E <= g,
    // so d1, d2 >= 1 - g >= 1e-6 (clamped gamma) and d1 d2 is a normal float.
    __asm__ volatile("" ::: "memory");
    float E = 0.f;
#pragma unroll
    for (int r = 0; r < kL; r++) {
        if ((!VARLEN || r < L) && r >= lcr) {
            float Tn;
            if (VARLEN) Tn = (r + 1 < L) ? pick10(V, L - 1 - r) : 1.f;
            else Tn = (r + 1 < kL) ? V[(kL - 1 - r) < 0 ? 0 : (kL - 1 - r)] : 1.f;
            if (r == lcr) {
                const float sv = src.get(1, pv[r]);
                // At most 1 (g car <= 1), so the unit conversion applies.
                em.add(1, r, pv[r], fdiv(sv, 1 - (1 - sv) * g * car));
                const float phi4 = (1 - sv) * (1 - g);
                const float phi5 = (1 - sv) * g * Tn;
                const float inv = __builtin_amdgcn_rcpf(phi4 + phi5 + sv);
                gn += phi5 * inv;
                gd += (phi4 + phi5) * inv;
                E = g * (1 - sv);
            } else {
                const float a = src.get(0, pv[r]);
                const float d1 = 1 - g * E * (1 - Tn);
                const float d2 = 1 - a * E;
                const float q = __builtin_amdgcn_rcpf(d1 * d2);
                const float w = E * (d2 * q);
                gn += g * Tn * w;
                gd += (1 - g + g * Tn) * w;
                E = E * (g * (1 - a)) * (d1 * q);
            }
        }
    }
}

// CCM E-step, ccm.cu:185-467, for a session with at least one click.
//
// * compute_exm_car (ccm.cu:245-279): t_r = (1-a_r) tau1,
//   exam[r+1] = exam[r](t_r + a_r((1-a_r) tau2 + a_r tau3)), car as DBN.
// * attractiveness (ccm.cu:295-326): clicked ranks give (1,2) plus, at the
//   last click, a/(1-(tau2(1-a)+tau3 a) car[r+1]); unclicked ranks after
//   the last click give (1-exam)a/(1-exam car).
// * taus (ccm.cu:387-467, factor.cu:72-142): unclicked r < lcr give (1,1)
//   to tau1 exactly (constant).  Clicked r < lcr keep phi5 and phi7 only;
//   the tail product and examination cancel, leaving
//   x = (1-a)tau2 / ((1-a)tau2 + a tau3) to tau2 as (x, x) and 1-x to tau3.
//   At the last click and after it the tail product of start r+1 is
//   V_{L-1-r} with G_{k+1} = G_k tau1 (1-a_k).  The reference forms the
//   denominators num + phi/sum in double and narrows them (ccm.cu:433-467);
//   the per-lane float sums here round each term once, which is within the
//   same tolerance.
template <bool VARLEN, class Src, class Emit>
__device__ __forceinline__ void ccm_session(const int (&pv)[kL], int L, uint32_t mm, const Src& src, Emit& em,
                                            float t1, float t2, float t3, float* ta) {
#pragma clang fp contract(fast)
    const int lcr = last_click(mm);
    // Pass 1, ranks 0..L-1: exam[r] (compute_exm_car) and the tail products
    // V_k of the fresh-list chain (G_{k+1} = G_k tau1 (1 - a_k)).
    // The examination factor (1-a) tau1 + a((1-a) tau2 + a tau3) is the
    // quadratic tau1 + a (tau2 - tau1) + a^2 (tau3 - tau2): two multiply-adds.
    float exam[kL], V[kL];
    {
        const float c1 = t2 - t1, c2 = t3 - t2;
        float e = 1.f, v = 1.f, G = 1.f;
#pragma unroll
        for (int r = 0; r < kL; r++) {
            exam[r] = e;
            V[r] = v;
            if (!VARLEN || r < L) {
                const float a = src.get(0, pv[r]);
                e = e * ((c2 * a + c1) * a + t1);
                v = v - a * G;
                G = G * (t1 - a * t1);
            }
        }
    }
    // Pass 2, ranks L-1 down to lcr+1 (all unclicked): car[r] and the
    // attractiveness posterior (1-exam)a/(1-exam car); car[lcr+1] is kept
    // for the last click.  Parameters are read again from LDS.
    __asm__ volatile("" ::: "memory");
    float car = 0.f;
#pragma unroll
    for (int r = kL - 1; r >= 0; r--) {
        if ((!VARLEN || r < L) && r > lcr) {
            const float a = src.get(0, pv[r]);
            car = a + (t1 - a * t1) * car;
            em.add(0, r, pv[r], fdiv(a - exam[r] * a, 1 - exam[r] * car));
        }
    }
    // Pass 3, ranks 0..L-1: tau2/tau3 posteriors of earlier clicks, the last
    // click, and the tau1 factors after it.
    __asm__ volatile("" ::: "memory");
    float ex = 0.f;
#pragma unroll
    for (int r = 0; r < kL; r++) {
        if (VARLEN && r >= L) continue;
        const bool clicked = (mm >> r) & 1u;
        if (r < lcr) {
            if (clicked) {
                const float a = src.get(0, pv[r]);
                const float w2 = (1 - a) * t2;
                const float w3 = a * t3;
                const float inv = __builtin_amdgcn_rcpf(w2 + w3);
                const float x = w2 * inv;
                const float y = w3 * inv;
                ta[2] += x; ta[3] += x;
                ta[4] += y; ta[5] += y;
            }
            continue;
        }
        const float a = src.get(0, pv[r]);
        float Tn;
        if (VARLEN) Tn = (r + 1 < L) ? pick10(V, L - 1 - r) : 1.f;
        else Tn = (r + 1 < kL) ? V[(kL - 1 - r) < 0 ? 0 : (kL - 1 - r)] : 1.f;
        if (r == lcr) {
            em.add_wide(0, r, pv[r], fdiv(a, 1 - (t2 * (1 - a) + t3 * a) * car));
            const float phi4 = (1 - a) * (1 - t2);
            const float phi5 = (1 - a) * t2 * Tn;
            const float phi6 = a * (1 - t3);
            const float phi7 = a * t3 * Tn;
            const float inv = __builtin_amdgcn_rcpf(phi4 + phi5 + phi6 + phi7);
            const float n2 = phi5 * inv;
            const float n3 = phi7 * inv;
            ta[2] += n2;
            ta[3] += n2 + phi4 * inv;
            ta[4] += n3;
            ta[5] += n3 + phi6 * inv;
            ex = t2 * (1 - a) + t3 * a;
        } else {
            // The factor sum d1 and the examination update's denominator d2
            // share one reciprocal: q = 1/(d1 d2), 1/d1 = d2 q, 1/d2 = d1 q.
            // Both are at least 1e-6 (clamped parameters), so d1 d2 is a
            // normal float.
            const float E = ex;
            const float phi0 = 1 - E;
            const float phi4 = (1 - t1) * E;
            const float phi5 = t1 * Tn * E;
            const float d1 = phi0 + phi4 + phi5;
            const float d2 = 1 - a * E;
            const float q = __builtin_amdgcn_rcpf(d1 * d2);
            const float inv = d2 * q;
            const float n1 = phi5 * inv;
            ta[0] += n1;
            ta[1] += n1 + phi4 * inv;
            ex = E * (t1 * (1 - a)) * (d1 * q);
        }
    }
}

// --------------------------------------------------------------------------
// Shared by the EM and the gradient kernels.
// --------------------------------------------------------------------------

// Number of per-qd tables, shared-table length and shared accumulators.
template <int MODEL> struct Shape;
template <> struct Shape<kPBM> { static constexpr int T = 1, K = 10, KACC = 10; };
template <> struct Shape<kUBM> { static constexpr int T = 1, K = 55, KACC = 55; };
template <> struct Shape<kDBN> { static constexpr int T = 2, K = 1, KACC = 2; };
template <> struct Shape<kCCM> { static constexpr int T = 1, K = 3, KACC = 6; };

// Buffer descriptors of the session stream: 32-bit offsets (no 64-bit
// address arithmetic per load) and hardware bounds checks (a prefetch past
// the last session reads zeros).
struct Stream {
    __amdgpu_buffer_rsrc_t packed, pidx16, pidx, mask, len;
    uint32_t row16_bytes, row_bytes, n_w, n_pk;
};

__device__ __forceinline__ Stream make_stream(const KArgs& A) {
    constexpr int kRawDwordFlags = 0x00020000;   // raw 32-bit buffer, no format conversion
    Stream st;
    const uint32_t n_rest = (uint32_t)(A.N - A.n_w);
    const uint32_t n_w16 = (uint32_t)(A.n_w - A.n_pk);
    st.packed = __builtin_amdgcn_make_buffer_rsrc((void*)A.packed, 0, (int)(uint32_t)(A.n_pk * 16), kRawDwordFlags);
    st.pidx16 = __builtin_amdgcn_make_buffer_rsrc((void*)A.pidx16, 0, (int)(n_w16 * 20u), kRawDwordFlags);
    st.pidx = __builtin_amdgcn_make_buffer_rsrc((void*)A.pidx, 0, (int)(n_rest * 40u), kRawDwordFlags);
    st.mask = __builtin_amdgcn_make_buffer_rsrc((void*)A.mask, 0, (int)(uint32_t)(A.N * 2), kRawDwordFlags);
    st.len = A.len ? __builtin_amdgcn_make_buffer_rsrc((void*)A.len, 0, (int)(uint32_t)A.N, kRawDwordFlags) : st.mask;
    st.row16_bytes = n_w16 * 2u;
    st.row_bytes = n_rest * 4u;
    st.n_w = (uint32_t)A.n_w;
    st.n_pk = (uint32_t)A.n_pk;
    return st;
}

// Index of rank r of the session at device position o: a 16-bit window
// offset for window items stored rank-major (WIN), else an absolute 32-bit
// index.
template <bool WIN>
__device__ __forceinline__ int load_index(const Stream& st, uint32_t o, int r) {
    if (WIN) return (int)__builtin_amdgcn_raw_buffer_load_b16(st.pidx16, (o - st.n_pk) * 2u, r * st.row16_bytes, 0);
    return (int)__builtin_amdgcn_raw_buffer_load_b32(st.pidx, (o - st.n_w) * 4u, r * st.row_bytes, 0);
}

// Loads one session's click bits, length and query-document indices
// (rank-major layout: rank r of the session at position s is element
// r * rows + s of its stream).
// Lanes past the end of their item (s >= s_end) re-read the item's last
// session, whose cache lines are already fetched: no branch, no extra DRAM
// traffic, and the values are never used.  Lanes without an item (s_end <= 0)
// compute a wrapped offset: whatever such a load returns is never used.
template <bool VARLEN, bool WIN>
__device__ __forceinline__ void load_session(const Stream& st, int s, int s_end, uint32_t& m, int& L, int (&pv)[kL]) {
    const uint32_t o = (uint32_t)min(s, s_end - 1);
    m = __builtin_amdgcn_raw_buffer_load_b16(st.mask, o * 2u, 0, 0);
    L = VARLEN ? (int)__builtin_amdgcn_raw_buffer_load_b8(st.len, o, 0, 0) : kL;
#pragma unroll
    for (int r = 0; r < kL; r++) pv[r] = load_index<WIN>(st, o, r);
}

// Click bits and length of one session (index clamped as in load_session).
__device__ __forceinline__ void load_head(const Stream& st, int s, int s_end, uint32_t& m, int& L) {
    const uint32_t o = (uint32_t)min(s, s_end - 1);
    m = __builtin_amdgcn_raw_buffer_load_b16(st.mask, o * 2u, 0, 0);
    L = (int)__builtin_amdgcn_raw_buffer_load_b8(st.len, o, 0, 0);
}

// Query-document indices of the first L ranks of one session; ranks past L
// are not read (a wavefront whose sessions all end before rank r issues no
// load for it).
template <bool WIN>
__device__ __forceinline__ void load_pidx_upto(const Stream& st, int s, int s_end, int L, int (&pv)[kL]) {
    const uint32_t o = (uint32_t)min(s, s_end - 1);
#pragma unroll
    for (int r = 0; r < kL; r++) {
        pv[r] = 0;
        if (r < L) pv[r] = load_index<WIN>(st, o, r);
    }
}

// A packed session record (16 bytes, one load per session):
//   bytes 0-9    offset of rank r's index from the session's base (ranks
//                past the length hold 0),
//   bytes 10-11  click bits,
//   bytes 12-13  base: the smallest of the session's window-relative indices,
//   byte  14     ranks shown, byte 15 unused.
typedef unsigned int PackedSession __attribute__((ext_vector_type(4)));

// The record of the session at position s (clamped as in load_session).
__device__ __forceinline__ PackedSession load_packed(const Stream& st, int s, int s_end) {
    const uint32_t o = (uint32_t)min(s, s_end - 1);
    return __builtin_amdgcn_raw_buffer_load_b128(st.packed, o * 16u, 0, 0);
}

// Click bits, length and window-relative indices of a packed record.
template <bool VARLEN>
__device__ __forceinline__ void decode_packed(PackedSession v, uint32_t& m, int& L, int (&pv)[kL]) {
    const int base = (int)(v[3] & 0xffffu);
    m = v[2] >> 16;
    L = VARLEN ? (int)((v[3] >> 16) & 0xffu) : kL;
#pragma unroll
    for (int r = 0; r < kL; r++) pv[r] = base + (int)((v[r >> 2] >> (8 * (r & 3))) & 0xffu);
}

// Wave-level sum of a 64-bit integer (wave64).
__device__ __forceinline__ long long wave_sum(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed point (32.32) plus a float base, rounded to float.
__device__ __forceinline__ float fix_to_float(float base, unsigned long long v) {
    return (float)((double)base + (double)(long long)v * kFixInv);
}

}  // namespace mc

#endif  // MC_KERNELS_H
