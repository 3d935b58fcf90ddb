// mc_layout.cpp -- host-side planning of a model's device layout, done once
// by mc_init_parameters: validation of the training sessions, the
// parameter-independent counts folded into the M-step, the work items
// (query groups packed under LDS windows, slices of large groups, direct
// single-session items), the device order of the sessions and their
// records and rank-major index streams, and the minibatch lists of the
// gradient path.
// The device allocation and upload stay in mc_engine.hip.

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"
#include "mc_layout.h"
#include "mc_model.h"

namespace mc {

namespace {
// Per-qd window: 1024 entries per per-qd table (an 8-byte accumulator and
// a staged 4-byte value each: 24 KB of LDS for the DBN) unless the model's
// MC_OPT_ITEM_WINDOW says otherwise.
int default_window(int T) { (void)T; return 1024; }
}  // namespace

mc_status plan_layout(mc_model* m, int64_t n, const int32_t* query, const int32_t* pidx, const uint8_t* clicks,
                      const uint8_t* len, int64_t n_qd, HostLayout* h) {
    const int T = m->T, K = m->K;
    // ---- validate and summarise the sessions ---------------------------------
    std::vector<uint8_t> L((size_t)n, (uint8_t)kMaxSerp);
    std::vector<uint16_t> mask((size_t)n, 0);
    std::vector<uint8_t> lcr((size_t)n, 0);
    for (int64_t s = 0; s < n; s++) {
        int l = kMaxSerp;
        if (len) {
            l = len[s];
            MC_REQUIRE(l >= 1 && l <= kMaxSerp, "len must be in [1, 10] (session " + std::to_string(s) + ")");
        }
        L[s] = (uint8_t)l;
        uint16_t mk = 0;
        for (int r = 0; r < l; r++) {
            const int32_t p = pidx[s * kMaxSerp + r];
            MC_REQUIRE(p >= 0 && p < n_qd, "pidx out of [0, n_qd) at session " + std::to_string(s));
            const uint8_t c = clicks[s * kMaxSerp + r];
            MC_REQUIRE(c <= 1, "clicks must be 0 or 1");
            if (c) mk |= (uint16_t)(1u << r);
        }
        mask[s] = mk;
        lcr[s] = mk ? (uint8_t)(31 - __builtin_clz((unsigned)mk)) : (uint8_t)kMaxSerp;
    }

    // ---- constants: per-qd (base numerator, denominator), shared bases -------
    // occ/clicked counts per entry; see the session functions in mc_kernels.h
    // for which counts are parameter-independent.
    std::vector<int64_t> occ((size_t)n_qd, 0), clk((size_t)n_qd, 0);
    for (int64_t s = 0; s < n; s++)
        for (int r = 0; r < L[s]; r++) {
            const int32_t p = pidx[s * kMaxSerp + r];
            occ[p]++;
            if ((mask[s] >> r) & 1u) clk[p]++;
        }
    std::vector<int2> counts((size_t)n_qd);
    for (int64_t p = 0; p < n_qd; p++) {
        MC_REQUIRE(occ[p] < INT32_MAX, "too many occurrences of one query-document pair");
        counts[p] = make_int2((int)occ[p], (int)clk[p]);
    }
    m->h_counts = counts;
    std::vector<double> sh_base(2 * (size_t)K, 0.0);
    std::vector<int32_t> sh_map(2 * (size_t)K, -1);
    switch (m->type) {
        case MC_PBM:
        case MC_UBM: {
            // w[k]: sessions whose temporary slot k is written (by the last
            // rank writing it, see ubm_session), wc[k]: of those, written by
            // a clicked rank with (1,1).  Unwritten slots keep (1,2).
            std::vector<int64_t> w((size_t)K, 0), wc((size_t)K, 0);
            for (int64_t s = 0; s < n; s++) {
                int last = kMaxSerp - 1;
                int slot[kMaxSerp];
                for (int r = 0; r < L[s]; r++) {
                    slot[r] = m->type == MC_PBM ? r : r * (r + 1) / 2 + last;
                    if ((mask[s] >> r) & 1u) last = r;
                }
                for (int r = 0; r < L[s]; r++) {
                    bool last_writer = true;
                    for (int q = r + 1; q < L[s]; q++) last_writer = last_writer && slot[q] != slot[r];
                    if (!last_writer) continue;
                    w[slot[r]]++;
                    if ((mask[s] >> r) & 1u) wc[slot[r]]++;
                }
            }
            for (int k = 0; k < K; k++) {
                const double ph = (double)(n - w[k]);
                sh_base[2 * k] = 1.0 + (double)wc[k] + ph;
                sh_base[2 * k + 1] = 2.0 + (double)w[k] + 2.0 * ph;
                sh_map[2 * k] = k;
            }
            break;
        }
        case MC_DBN: {
            int64_t n0 = 0;
            for (int64_t s = 0; s < n; s++) n0 += std::min<int64_t>(lcr[s], L[s]);
            sh_base[0] = 1.0 + (double)n0;
            sh_base[1] = 2.0 + (double)n0;
            sh_map[0] = 0;
            sh_map[1] = 1;
            break;
        }
        default: {
            int64_t n1 = 0;
            for (int64_t s = 0; s < n; s++)
                for (int r = 0; r < std::min<int>(lcr[s], L[s]); r++)
                    if (!((mask[s] >> r) & 1u)) n1++;
            sh_base[0] = 1.0 + (double)n1;
            sh_base[1] = 2.0 + (double)n1;
            for (int k = 1; k < 3; k++) { sh_base[2 * k] = 1.0; sh_base[2 * k + 1] = 2.0; }
            for (int k = 0; k < 3; k++) { sh_map[2 * k] = 2 * k; sh_map[2 * k + 1] = 2 * k + 1; }
            break;
        }
    }

    // ---- work items -----------------------------------------------------------
    // An item is a set of whole query groups, or a slice of one large group,
    // whose sessions are laid out contiguously in device order.  Kinds:
    //   window  groups packed under an entry window of at most W entries (LDS);
    //   global  slices of a group whose entries exceed W (global atomics);
    //   direct  single-session groups with distinct documents: each session
    //           owns its entries, so no reduction is needed at all.
    const int W = m->item_window > 0 ? m->item_window : default_window(T);
    const int S = m->item_sessions;
    enum { kKindWindow = 0, kKindGlobal = 1, kKindDirect = 2 };
    struct Spec {
        std::vector<std::pair<int64_t, int64_t>> ranges;
        int64_t count = 0;
        int32_t base = 0, width = 0, flags = 0;
        int kind = kKindWindow;
        bool packed = false;    // window item stored as 16-byte session records
        int32_t split = -1;     // slice of split group `split` (window slices only)
        int64_t gstart = 0;     // first session (query order) of the item's first group
    };
    std::vector<int4> split_groups;   // per window-sliced group: first entry, end entry, slices
    std::vector<Spec> specs;
    std::vector<int32_t> shared_idx, shared_batch;
    // Minibatches of the gradient path: contiguous runs of whole query
    // groups, cut by session count.
    const int B = m->n_batches;
    auto batch_of = [&](int64_t first_session) { return (int)std::min<int64_t>(B - 1, first_session * B / n); };
    // Both reduction modes use query-group work items when the training set
    // is grouped by query; the float-atomic mode merges every entry.
    bool structured = true;
    const bool allow_direct = m->direct && m->reduce_mode == MC_REDUCE_SEGMENTED;
    if (structured) {
        // Query groups: runs of equal query ids.  The segmented reduction needs
        // each group's entries in a range [lo, hi) disjoint from and after the
        // previous group's, which first-occurrence indexing of a query-sorted
        // train set guarantees (dataset.cpp:314-339).
        struct Group { int64_t s0, s1; int32_t lo, hi; };
        std::vector<Group> groups;
        int64_t prev_hi = 0;
        for (int64_t s = 0; s < n && structured;) {
            int64_t e = s;
            int32_t lo = INT32_MAX, hi = -1;
            while (e < n && query[e] == query[s]) {
                for (int r = 0; r < L[e]; r++) {
                    lo = std::min(lo, pidx[e * kMaxSerp + r]);
                    hi = std::max(hi, pidx[e * kMaxSerp + r] + 1);
                }
                e++;
            }
            if (lo < prev_hi) structured = false;
            prev_hi = hi;
            groups.push_back(Group{s, e, lo, hi});
            s = e;
        }
        if (structured) {
            auto distinct = [&](int64_t s) {
                for (int r = 0; r < L[s]; r++)
                    for (int q = r + 1; q < L[s]; q++)
                        if (pidx[s * kMaxSerp + r] == pidx[s * kMaxSerp + q]) return false;
                return true;
            };
            // Direct items are short (two sessions per lane): there are few
            // such sessions and each is independent, so many small
            // workgroups spread them over the whole chip.
            const int64_t S_dir = 2 * (int64_t)m->block;
            Spec cur, dir;
            bool open = false;
            auto close = [&]() { if (open) { specs.push_back(cur); open = false; } };
            auto close_dir = [&]() {
                if (dir.count) { dir.kind = kKindDirect; dir.flags = kItemDirect; specs.push_back(dir); dir = Spec(); }
            };
            for (const Group& g : groups) {
                const int64_t gs = g.s1 - g.s0;
                const int64_t gw = g.hi - g.lo;
                if (allow_direct && gs == 1 && distinct(g.s0)) {
                    if (dir.count == 0) dir.gstart = g.s0;
                    dir.ranges.push_back({g.s0, g.s1});
                    if (++dir.count == S_dir) close_dir();
                    continue;
                }
                if (gs <= S && gw <= W) {
                    if (open && cur.count + gs <= S && (g.hi - cur.base) <= W &&
                        batch_of(g.s0) == batch_of(cur.gstart)) {
                        cur.ranges.push_back({g.s0, g.s1});
                        cur.count += gs;
                        cur.width = g.hi - cur.base;
                    } else {
                        close();
                        cur = Spec();
                        cur.ranges.push_back({g.s0, g.s1});
                        cur.count = gs;
                        cur.base = g.lo;
                        cur.width = (int32_t)gw;
                        cur.flags = kItemInterior;
                        cur.gstart = g.s0;
                        open = true;
                    }
                } else {
                    close();
                    const int64_t n_slices = (gs + S - 1) / S;
                    const int32_t split = gw <= W ? (int32_t)split_groups.size() : -1;
                    if (gw <= W) split_groups.push_back(make_int4(g.lo, g.hi, (int)n_slices, 0));
                    for (int64_t k = 0; k < n_slices; k++) {
                        Spec sl;
                        const int64_t a = g.s0 + gs * k / n_slices, b = g.s0 + gs * (k + 1) / n_slices;
                        sl.ranges.push_back({a, b});
                        sl.count = b - a;
                        sl.base = g.lo;
                        sl.width = gw <= W ? (int32_t)gw : 0;
                        sl.kind = gw <= W ? kKindWindow : kKindGlobal;
                        sl.split = split;
                        sl.gstart = g.s0;
                        specs.push_back(sl);
                    }
                    for (int32_t p = g.lo; p < g.hi; p++) {
                        shared_idx.push_back(p);
                        shared_batch.push_back(batch_of(g.s0));
                    }
                }
            }
            close();
            close_dir();
        }
    }
    if (!structured) {
        // Float atomics over uniform chunks; the merge pass finalises every entry.
        m->reduce_mode = MC_REDUCE_ATOMIC;
        specs.clear();
        shared_idx.clear();
        split_groups.clear();
        for (int64_t a = 0; a < n; a += S) {
            Spec c;
            c.ranges.push_back({a, std::min<int64_t>(n, a + S)});
            c.count = c.ranges[0].second - a;
            c.kind = kKindGlobal;
            c.gstart = a;
            specs.push_back(c);
        }
    }
    // Packed window items: every session's indices lie within 256
    // consecutive entries of the window (a query's entries are numbered
    // contiguously, so this holds unless a query has more than 256 entries),
    // and the session is stored as one 16-byte record: 8-bit offsets from its
    // smallest index (segmented mode; the float-atomic kernels read the
    // rank-major stream).
    if (structured && m->packed && m->reduce_mode == MC_REDUCE_SEGMENTED) {
        for (Spec& sp : specs) {
            if (sp.kind != kKindWindow || sp.width > 65536) continue;
            bool fits = true;
            for (const auto& rg : sp.ranges)
                for (int64_t s = rg.first; s < rg.second && fits; s++) {
                    int32_t lo = INT32_MAX, hi = INT32_MIN;
                    for (int r = 0; r < L[s]; r++) {
                        lo = std::min(lo, pidx[s * kMaxSerp + r]);
                        hi = std::max(hi, pidx[s * kMaxSerp + r]);
                    }
                    fits = hi - lo <= 255;
                }
            sp.packed = fits;
            if (fits) sp.flags |= kItemPacked;
        }
    }
    // Launch order: packed window items, other window items, global items,
    // direct items; the largest items of each kind first (workgroups are
    // dispatched in order, so long items do not end up in the tail).
    std::stable_sort(specs.begin(), specs.end(), [](const Spec& a, const Spec& b) {
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.packed != b.packed) return a.packed;
        return a.count > b.count;
    });

    // Device order: each item's sessions contiguous, sorted by length and
    // last-click rank inside the item (wavefront-uniform chain lengths; the
    // per-entry sums do not depend on the order).  DBN and CCM sessions
    // without a click go last: by the reference's last-click convention
    // (search.cu:160-173 returns MAX_SERP when nothing is clicked) every rank
    // of such a session is "before the last click", so all its counts are the
    // constants already folded into the bases above, and the EM kernels stop
    // at the item's first such session (Item::e1).  The gradient path reads
    // them (their click-free log-likelihood depends on the parameters).
    // Direct items write their entries' values per session and visit all.
    const bool constant_clickless = m->type == MC_DBN || m->type == MC_CCM;
    std::vector<Item> items;
    std::vector<int64_t> order;
    order.reserve((size_t)n);
    m->n_win_items = m->n_pk_items = m->n_glob_items = m->n_dir_items = 0;
    m->n_em_sessions = 0;
    for (const Spec& sp : specs) {
        const int64_t j0 = (int64_t)order.size();
        for (const auto& rg : sp.ranges)
            for (int64_t s = rg.first; s < rg.second; s++) order.push_back(s);
        auto clickless = [&](int64_t s) { return constant_clickless && mask[s] == 0; };
        if (m->sort_by_lcr)
            std::stable_sort(order.begin() + j0, order.end(), [&](int64_t a, int64_t b) {
                if (clickless(a) != clickless(b)) return clickless(b);
                return L[a] != L[b] ? L[a] < L[b] : lcr[a] < lcr[b];
            });
        else
            std::stable_partition(order.begin() + j0, order.end(), [&](int64_t s) { return !clickless(s); });
        int64_t e1 = (int64_t)order.size();
        if (m->skip_constant && sp.kind != kKindDirect)
            while (e1 > j0 && clickless(order[(size_t)e1 - 1])) e1--;
        m->n_em_sessions += e1 - j0;
        items.push_back(Item{(int32_t)j0, (int32_t)order.size(), sp.base, sp.width, sp.flags, batch_of(sp.gstart),
                             (int32_t)e1, sp.split});
        if (sp.kind == kKindWindow && structured) {
            m->n_win_items++;
            if (sp.packed) m->n_pk_items++;
        }
        else if (sp.kind == kKindDirect) m->n_dir_items++;
        else m->n_glob_items++;
    }
    MC_REQUIRE((int64_t)order.size() == n, "internal: work items do not cover the sessions");
    // Every session of a window item must index inside that window (the
    // kernel addresses LDS with it); items must tile [0, n).
    {
        int64_t expect = 0;
        for (const Item& it : items) {
            MC_REQUIRE(it.s0 == expect && it.s1 > it.s0 && it.s1 <= n && it.e1 >= it.s0 && it.e1 <= it.s1,
                       "internal: work items do not tile the sessions");
            expect = it.s1;
            if (it.width > 0) {
                MC_REQUIRE(it.base >= 0 && (int64_t)it.base + it.width <= n_qd && it.width <= W,
                           "internal: item window out of range");
                for (int64_t j = it.s0; j < it.s1; j++) {
                    const int64_t s = order[j];
                    for (int r = 0; r < L[s]; r++) {
                        const int32_t p = pidx[s * kMaxSerp + r];
                        MC_REQUIRE(p >= it.base && p < it.base + it.width, "internal: session outside its item window");
                    }
                }
            }
        }
        MC_REQUIRE(expect == n, "internal: work items do not cover the sessions");
    }
    m->n_items = (int)items.size();
    m->n_shared = (int64_t)shared_idx.size();

    // Minibatch lists of the gradient path: per batch its window items and
    // its global items (item indices), then the same for every item; split
    // entries ordered by batch.
    std::vector<int32_t> lists;
    {
        std::vector<size_t> perm(shared_idx.size());
        std::iota(perm.begin(), perm.end(), (size_t)0);
        std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return shared_batch[a] < shared_batch[b]; });
        std::vector<int32_t> sorted(shared_idx.size());
        for (size_t i = 0; i < perm.size(); i++) sorted[i] = shared_idx[perm[i]];
        m->shared_off.assign((size_t)B + 1, 0);
        m->shared_cnt.assign((size_t)B + 1, 0);
        for (size_t i = 0; i < perm.size(); i++) m->shared_cnt[(size_t)shared_batch[perm[i]]]++;
        for (int b = 1; b < B; b++) m->shared_off[(size_t)b] = m->shared_off[(size_t)b - 1] + m->shared_cnt[(size_t)b - 1];
        m->shared_off[(size_t)B] = 0;
        m->shared_cnt[(size_t)B] = (int64_t)sorted.size();
        shared_idx.swap(sorted);
        m->win_off.assign((size_t)B + 1, 0);
        m->win_cnt.assign((size_t)B + 1, 0);
        m->glob_off.assign((size_t)B + 1, 0);
        m->glob_cnt.assign((size_t)B + 1, 0);
        m->batch_sessions.assign((size_t)B, 0);
        for (const Item& it : items) m->batch_sessions[(size_t)it.pad] += it.s1 - it.s0;
        const int g0 = m->n_win_items, g1 = m->n_win_items + m->n_glob_items;
        for (int b = 0; b <= B; b++) {
            m->win_off[(size_t)b] = (int64_t)lists.size();
            for (int i = 0; i < g0; i++)
                if (b == B || items[(size_t)i].pad == b) lists.push_back(i);
            m->win_cnt[(size_t)b] = (int64_t)lists.size() - m->win_off[(size_t)b];
            m->glob_off[(size_t)b] = (int64_t)lists.size();
            for (int i = g0; i < g1; i++)
                if (b == B || items[(size_t)i].pad == b) lists.push_back(i);
            m->glob_cnt[(size_t)b] = (int64_t)lists.size() - m->glob_off[(size_t)b];
        }
    }
    m->max_item_sessions = 0;
    for (const Item& it : items) m->max_item_sessions = std::max<int64_t>(m->max_item_sessions, it.s1 - it.s0);
    int max_width = 0;
    for (const Item& it : items) max_width = std::max(max_width, (int)it.width);
    m->max_width = max_width;
    m->lds_bytes = (size_t)max_width * T * (m->reduce_mode == MC_REDUCE_SEGMENTED ? 12 : 8);
    // The UBM's fixed-point windows are followed by its examination slots.
    if (m->type == MC_UBM && m->reduce_mode == MC_REDUCE_SEGMENTED)
        m->lds_bytes = (m->lds_bytes + 7) / 8 * 8 + kUbmSlotBytes;
    // ---- device layout: session records, rank-major pidx, clicks, lengths --
    // Window items come first in device order, packed ones first of all.
    // Their sessions store indices relative to the window base (the kernels
    // address the LDS window with them): 16-byte records (8-bit offsets from
    // the session's smallest index, which is stored with the click bits and
    // the length), else 16-bit indices, 20 bytes per session instead of 40.
    // The other sessions store absolute 32-bit indices.  Click bits and
    // lengths are also kept per position for every session.
    const int64_t n_w = m->n_win_items > 0 ? (int64_t)items[(size_t)m->n_win_items - 1].s1 : 0;
    const int64_t n_pk = m->n_pk_items > 0 ? (int64_t)items[(size_t)m->n_pk_items - 1].s1 : 0;
    const int64_t n_w16 = n_w - n_pk;
    const int64_t n_rest = n - n_w;
    MC_REQUIRE(n_rest * 40 < (int64_t)UINT32_MAX,
               "too many sessions outside window items (" + std::to_string(n_rest) +
                   ", at most 107,374,182 with 32-bit indices): use a larger MC_OPT_ITEM_WINDOW or fewer sessions");
    MC_REQUIRE(n_pk * 16 < (int64_t)UINT32_MAX && n_w16 * 20 < (int64_t)UINT32_MAX, "internal: session stream size");
    m->n_w = n_w;
    m->n_pk = n_pk;
    std::vector<uint32_t> h_packed((size_t)n_pk * 4);
    std::vector<uint16_t> h_pidx16((size_t)n_w16 * kMaxSerp);
    std::vector<int32_t> h_pidx((size_t)n_rest * kMaxSerp);
    std::vector<uint16_t> h_mask((size_t)n);
    std::vector<uint8_t> h_len((size_t)n);
    for (size_t i = 0; i < items.size(); i++) {
        const Item& it = items[i];
        const bool win = (int)i < m->n_win_items;
        const bool pk = (int)i < m->n_pk_items;
        MC_REQUIRE(!win || (it.width > 0 && it.width <= 65536 && it.s1 <= n_w), "internal: window item layout");
        MC_REQUIRE(win || it.s0 >= n_w, "internal: window items must come first");
        MC_REQUIRE(pk == ((it.flags & kItemPacked) != 0) && (!pk || it.s1 <= n_pk) && (pk || it.s0 >= n_pk),
                   "internal: packed items must come first");
        for (int64_t j = it.s0; j < it.s1; j++) {
            const int64_t s = order[j];
            if (pk) {
                int32_t lo = INT32_MAX;
                for (int r = 0; r < L[s]; r++) lo = std::min(lo, pidx[s * kMaxSerp + r] - it.base);
                uint32_t w[4] = {0u, 0u, (uint32_t)mask[s] << 16, (uint32_t)lo | ((uint32_t)L[s] << 16)};
                for (int r = 0; r < L[s]; r++) {
                    const int32_t off = pidx[s * kMaxSerp + r] - it.base - lo;
                    MC_REQUIRE(off >= 0 && off <= 255 && lo >= 0 && lo <= 65535, "internal: packed session range");
                    w[r >> 2] |= (uint32_t)off << (8 * (r & 3));
                }
                for (int k = 0; k < 4; k++) h_packed[(size_t)j * 4 + k] = w[k];
            } else {
                for (int r = 0; r < kMaxSerp; r++) {
                    const int32_t p = r < L[s] ? pidx[s * kMaxSerp + r] : 0;
                    if (win) h_pidx16[(size_t)r * n_w16 + (j - n_pk)] = r < L[s] ? (uint16_t)(p - it.base) : 0;
                    else h_pidx[(size_t)r * n_rest + (j - n_w)] = p;
                }
            }
            h_mask[j] = mask[s];
            h_len[j] = L[s];
        }
    }


    h->counts = std::move(counts);
    h->sh_base = std::move(sh_base);
    h->sh_map = std::move(sh_map);
    h->items = std::move(items);
    h->shared_idx = std::move(shared_idx);
    h->lists = std::move(lists);
    h->split_groups = std::move(split_groups);
    h->packed = std::move(h_packed);
    h->pidx16 = std::move(h_pidx16);
    h->pidx32 = std::move(h_pidx);
    h->mask = std::move(h_mask);
    h->len = std::move(h_len);
    return MC_OK;
}

}  // namespace mc
