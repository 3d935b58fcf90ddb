// mc_dataset.cpp -- host-side dataset preparation of the EM engine.
//
// Turns a raw click log (or in-memory sessions) into the training and test
// sets the reference builds in Dataset::parse_dataset (data/dataset.cpp:421-455):
//   * parsing: query lines of 15 tokens open a SERP, click lines of 4 tokens
//     mark the first rank showing the clicked document, a SERP is emitted only
//     when the next query line arrives (dataset.cpp:224-305, search.cu:93-122);
//   * split: n_train = int(n * (1 - test_share) - 0.5) + 1, the tail is the test
//     set (dataset.cpp:385-396);
//   * the train set is ordered by std::sort on the query id
//     (dataset.cpp:17-21); std::sort is not stable, so the same library
//     algorithm is used here to reproduce the reference's order of equal keys;
//   * query-document indices in first-occurrence order over the sorted train
//     set (dataset.cpp:314-339); test sessions with an unseen query or no seen
//     document are dropped, unseen documents keep index -1 (dataset.cpp:349-374).
//
// Two undefined behaviours of the reference parser get a defined meaning
// here: empty lines are skipped, and lines before the first query line never
// emit a SERP (the reference would emit an uninitialised one).

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "mc_internal.h"

namespace mc {

namespace {

constexpr int kQueryTokens = 15;   // utils/definitions.h:51
constexpr int kClickTokens = 4;    // utils/definitions.h:52

inline bool is_space(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

enum IntParse { kIntOk = 0, kIntInvalid = 1, kIntRange = 2 };

// std::stoi on one whitespace-free token: optional sign, decimal digits, stops
// at the first non-digit; no digits -> invalid_argument; outside int -> out_of_range.
IntParse parse_int(const char* b, const char* e, int* out) {
    const char* p = b;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) { neg = *p == '-'; p++; }
    if (p >= e || *p < '0' || *p > '9') return kIntInvalid;
    long long v = 0;
    bool overflow = false;
    for (; p < e && *p >= '0' && *p <= '9'; p++) {
        if (!overflow) {
            v = v * 10 + (*p - '0');
            if (v > (long long)INT_MAX + 1) overflow = true;
        }
    }
    if (neg) v = -v;
    if (overflow || v > INT_MAX || v < INT_MIN) return kIntRange;
    *out = (int)v;
    return kIntOk;
}

// Open-addressing int -> int map with O(1) clear through generation stamps;
// used to number the documents of one query group.
class GroupDocMap {
public:
    void reset(size_t expected) {
        size_t want = 16;
        while (want < expected * 2) want <<= 1;
        if (want > keys_.size()) {
            keys_.assign(want, 0);
            vals_.assign(want, 0);
            gen_.assign(want, 0);
            cur_ = 0;
        }
        cur_++;
        if (cur_ == 0) { std::fill(gen_.begin(), gen_.end(), 0u); cur_ = 1; }
        size_ = 0;
    }
    // Returns the stored index for key, inserting next_index if absent.
    int find_or_insert(int key, int next_index, bool* inserted) {
        if ((size_ + 1) * 2 > keys_.size()) grow();
        size_t mask = keys_.size() - 1;
        size_t h = hash(key) & mask;
        while (gen_[h] == cur_) {
            if (keys_[h] == key) { *inserted = false; return vals_[h]; }
            h = (h + 1) & mask;
        }
        gen_[h] = cur_;
        keys_[h] = key;
        vals_[h] = next_index;
        size_++;
        *inserted = true;
        return next_index;
    }

private:
    static size_t hash(int k) { return (size_t)((uint32_t)k * 2654435761u) ^ ((uint32_t)k >> 16); }
    void grow() {
        std::vector<int> k2, v2;
        for (size_t i = 0; i < keys_.size(); i++)
            if (gen_[i] == cur_) { k2.push_back(keys_[i]); v2.push_back(vals_[i]); }
        size_t n = keys_.size() * 2;
        keys_.assign(n, 0);
        vals_.assign(n, 0);
        gen_.assign(n, 0);
        cur_ = 1;
        size_ = 0;
        bool ins;
        for (size_t i = 0; i < k2.size(); i++) find_or_insert(k2[i], v2[i], &ins);
    }
    std::vector<int> keys_, vals_;
    std::vector<uint32_t> gen_;
    uint32_t cur_ = 0;
    size_t size_ = 0;
};

// Runs f(a, b) over [0, n) in blocks of at least `grain` items on up to 16
// threads (the calling thread takes the first block).
template <class F>
void parallel_ranges(int64_t n, int64_t grain, F f) {
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hw, 16), n / std::max<int64_t>(1, grain)));
    if (nt <= 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t t = 1; t < nt; t++) th.emplace_back(f, n * t / nt, n * (t + 1) / nt);
    f((int64_t)0, n / nt);
    for (auto& x : th) x.join();
}

}  // namespace

// Split, sort and index `all` (log order) into ds.  Returns false with an
// error message when the split is empty.  len_all (optional, log order)
// gives the ranks each session shows; ranks >= len hold no result: they are
// not indexed (pidx -1) and do not count as seen documents of a test session.
bool prepare_dataset(HostDataset& ds, const std::vector<int32_t>& q_all, const std::vector<int32_t>& d_all,
                     const std::vector<uint8_t>& c_all, float test_share, std::string* err,
                     const std::vector<uint8_t>* len_all) {
    const int n = (int)q_all.size();
    ds.n_parsed = n;
    auto shown = [&](int src, int r) { return !len_all || r < (*len_all)[(size_t)src]; };
    if (n < 1) { *err = "no complete session"; return false; }
    const int n_train = static_cast<int>(n * (1 - test_share) - 0.5) + 1;   // dataset.cpp:390
    if (n_train < 1 || n_train > n) {
        *err = "test share leaves an invalid train/test split";
        return false;
    }
    // Order of the train set: libstdc++ std::sort on the query key
    // (dataset.cpp:19-20).  The permutation depends only on the comparison
    // outcomes, so sorting (query, position) pairs reproduces it exactly.
    struct Key { int32_t q; int32_t pos; };
    std::vector<Key> keys((size_t)n_train);
    for (int i = 0; i < n_train; i++) keys[i] = Key{q_all[i], i};
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) { return a.q < b.q; });

    ds.train_query.resize((size_t)n_train);
    ds.train_docs.resize((size_t)n_train * kMaxSerp);
    ds.train_clicks.resize((size_t)n_train * kMaxSerp);
    ds.train_pidx.resize((size_t)n_train * kMaxSerp);
    ds.train_len.clear();
    ds.test_len.clear();
    if (len_all) ds.train_len.resize((size_t)n_train);
    // Gather in sorted order (random reads: split over threads).
    parallel_ranges(n_train, 1 << 16, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++) {
            const int src = keys[(size_t)i].pos;
            ds.train_query[(size_t)i] = q_all[(size_t)src];
            std::memcpy(&ds.train_docs[(size_t)i * kMaxSerp], &d_all[(size_t)src * kMaxSerp], kMaxSerp * 4);
            std::memcpy(&ds.train_clicks[(size_t)i * kMaxSerp], &c_all[(size_t)src * kMaxSerp], kMaxSerp);
            if (len_all) ds.train_len[(size_t)i] = (*len_all)[(size_t)src];
        }
    });
    auto train_shown = [&](int i, int r) { return !len_all || r < ds.train_len[(size_t)i]; };

    // First-occurrence indices; equal queries are contiguous after the sort,
    // so each query group is numbered with a group-local map.
    // A query group's new indices are consecutive, so group g owns the
    // entries [entry_lo[g], entry_lo[g + 1]); entry_doc records each
    // entry's document for the test-set lookup below.
    GroupDocMap gmap;
    int next = 0;
    std::vector<int32_t> group_query, entry_lo, entry_doc;
    for (int s = 0; s < n_train;) {
        int e = s;
        while (e < n_train && ds.train_query[e] == ds.train_query[s]) e++;
        group_query.push_back(ds.train_query[s]);
        entry_lo.push_back(next);
        gmap.reset((size_t)std::min(e - s, 64) * kMaxSerp);
        for (int i = s; i < e; i++) {
            for (int r = 0; r < kMaxSerp; r++) {
                if (!train_shown(i, r)) {
                    ds.train_pidx[(size_t)i * kMaxSerp + r] = -1;
                    continue;
                }
                bool ins;
                const int d = ds.train_docs[(size_t)i * kMaxSerp + r];
                int idx = gmap.find_or_insert(d, next, &ins);
                if (ins) {
                    entry_doc.push_back(d);
                    next++;
                }
                ds.train_pidx[(size_t)i * kMaxSerp + r] = idx;
            }
        }
        s = e;
    }
    entry_lo.push_back(next);
    ds.n_qd = next;

    // Test set: keep log order, drop sessions whose query is unseen or whose
    // documents are all unseen.  A (query, doc) pair is looked up as the
    // query's group (binary search; group_query is ascending) and the doc
    // among the group's entries: a scan for small groups, a binary search in
    // a per-group sorted copy for large ones.
    constexpr int kScan = 32;
    const int n_groups = (int)group_query.size();
    std::vector<int64_t> sorted_at((size_t)n_groups + 1, 0);
    for (int g = 0; g < n_groups; g++) {
        const int w = entry_lo[(size_t)g + 1] - entry_lo[(size_t)g];
        sorted_at[(size_t)g + 1] = sorted_at[(size_t)g] + (w > kScan ? w : 0);
    }
    std::vector<std::pair<int32_t, int32_t>> sorted_docs((size_t)sorted_at[(size_t)n_groups]);
    parallel_ranges(n_groups, 4096, [&](int64_t a, int64_t b) {
        for (int64_t g = a; g < b; g++) {
            const int lo = entry_lo[(size_t)g], hi = entry_lo[(size_t)g + 1];
            if (hi - lo <= kScan) continue;
            auto* out = &sorted_docs[(size_t)sorted_at[(size_t)g]];
            for (int k = lo; k < hi; k++) out[k - lo] = {entry_doc[(size_t)k], k};
            std::sort(out, out + (hi - lo));
        }
    });
    auto find_pair = [&](int g, int d) -> int {
        const int lo = entry_lo[(size_t)g], hi = entry_lo[(size_t)g + 1];
        if (hi - lo <= kScan) {
            for (int k = lo; k < hi; k++)
                if (entry_doc[(size_t)k] == d) return k;
            return -1;
        }
        const auto* b = &sorted_docs[(size_t)sorted_at[(size_t)g]];
        const auto* e = b + (hi - lo);
        const auto* it = std::lower_bound(b, e, std::make_pair(d, INT32_MIN));
        return (it != e && it->first == d) ? it->second : -1;
    };
    // Lookups in parallel over blocks of the test range, then the kept
    // sessions are appended in log order.
    const int n_tail = n - n_train;
    std::vector<int32_t> t_pid((size_t)n_tail * kMaxSerp);
    std::vector<uint8_t> t_keep((size_t)n_tail, 0);
    parallel_ranges(n_tail, 1 << 14, [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; k++) {
            const int i = n_train + (int)k;
            const int q = q_all[(size_t)i];
            const auto git = std::lower_bound(group_query.begin(), group_query.end(), q);
            if (git == group_query.end() || *git != q) continue;
            const int g = (int)(git - group_query.begin());
            int32_t* pid = &t_pid[(size_t)k * kMaxSerp];
            bool found = false;
            for (int r = 0; r < kMaxSerp; r++) {
                pid[r] = shown(i, r) ? find_pair(g, d_all[(size_t)i * kMaxSerp + r]) : -1;
                if (pid[r] != -1) found = true;
            }
            t_keep[(size_t)k] = found;
        }
    });
    ds.test_query.clear();
    ds.test_docs.clear();
    ds.test_clicks.clear();
    ds.test_pidx.clear();
    size_t n_keep = 0;
    for (int k = 0; k < n_tail; k++) n_keep += t_keep[(size_t)k];
    ds.test_query.reserve(n_keep);
    ds.test_docs.reserve(n_keep * kMaxSerp);
    ds.test_clicks.reserve(n_keep * kMaxSerp);
    ds.test_pidx.reserve(n_keep * kMaxSerp);
    if (len_all) ds.test_len.reserve(n_keep);
    for (int k = 0; k < n_tail; k++) {
        if (!t_keep[(size_t)k]) continue;
        const size_t i = (size_t)n_train + (size_t)k;
        if (len_all) ds.test_len.push_back((*len_all)[i]);
        ds.test_query.push_back(q_all[i]);
        ds.test_docs.insert(ds.test_docs.end(), &d_all[i * kMaxSerp], &d_all[i * kMaxSerp] + kMaxSerp);
        ds.test_clicks.insert(ds.test_clicks.end(), &c_all[i * kMaxSerp], &c_all[i * kMaxSerp] + kMaxSerp);
        ds.test_pidx.insert(ds.test_pidx.end(), &t_pid[(size_t)k * kMaxSerp], &t_pid[(size_t)k * kMaxSerp] + kMaxSerp);
    }
    return true;
}

namespace {

// One parsed line of the log.  Lines that do not start with an integer
// (and empty lines) produce no event; any other line updates the previous
// session id (dataset.cpp:288).
struct LineEvent {
    enum Kind : uint8_t { kQuery = 0, kClick = 1, kOther = 2, kError = 3 };
    uint8_t kind;
    int32_t sid;
    int32_t value;      // query id (kQuery), clicked document (kClick), error code (kError)
    int64_t line;       // line number within the chunk (for errors)
    int64_t docs;       // index of the 10 documents in the chunk's doc list (kQuery)
};

enum { kErrSid = 1, kErrQuery = 2, kErrDoc = 3, kErrClick = 4 };

struct ChunkEvents {
    std::vector<LineEvent> ev;
    std::vector<int32_t> docs;
    int64_t lines = 0;
};

// Tokenises lines [p, end) (std::istringstream >> std::string per line) into
// events; stops at the first malformed id, which the in-order pass reports
// if the read gets that far.
void tokenize_chunk(const char* p, const char* end, ChunkEvents* out) {
    const char* tb[kQueryTokens + 1];
    const char* te[kQueryTokens + 1];
    int64_t line = 0;
    // Query lines carry 15 tokens (at least ~30 bytes), click lines 4.
    out->ev.reserve((size_t)(end - p) / 32 + 16);
    out->docs.reserve((size_t)(end - p) / 12 + 16);
    while (p < end) {
        const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
        const char* le = nl ? nl : end;
        line++;
        int nt = 0;
        const char* t = p;
        while (t < le) {
            while (t < le && is_space(*t)) t++;
            if (t >= le) break;
            const char* s0 = t;
            while (t < le && !is_space(*t)) t++;
            if (nt <= kQueryTokens) { tb[nt] = s0; te[nt] = t; }
            nt++;
        }
        p = nl ? nl + 1 : end;
        if (nt == 0) continue;
        int sid;
        const IntParse ps = parse_int(tb[0], te[0], &sid);
        if (ps == kIntInvalid) continue;                       // dataset.cpp:258-263
        LineEvent e{LineEvent::kOther, sid, 0, line, -1};
        if (ps == kIntRange) {
            e.kind = LineEvent::kError;
            e.value = kErrSid;
            out->ev.push_back(e);
            break;
        }
        if (nt == kQueryTokens) {
            int q;
            if (parse_int(tb[3], te[3], &q) != kIntOk) {
                e.kind = LineEvent::kError;
                e.value = kErrQuery;
                out->ev.push_back(e);
                break;
            }
            const int64_t at = (int64_t)out->docs.size();
            bool ok = true;
            for (int r = 0; r < kMaxSerp && ok; r++) {
                int d;
                ok = parse_int(tb[r + 5], te[r + 5], &d) == kIntOk;
                out->docs.push_back(d);
            }
            if (!ok) {
                e.kind = LineEvent::kError;
                e.value = kErrDoc;
                out->ev.push_back(e);
                break;
            }
            e.kind = LineEvent::kQuery;
            e.value = q;
            e.docs = at;
        } else if (nt == kClickTokens) {
            int d;
            if (parse_int(tb[3], te[3], &d) != kIntOk) {
                e.kind = LineEvent::kError;
                e.value = kErrClick;
                out->ev.push_back(e);
                break;
            }
            e.kind = LineEvent::kClick;
            e.value = d;
        }
        out->ev.push_back(e);
    }
    out->lines = line;
}

}  // namespace

// Reads the raw log (dataset.cpp:224-305): lines are tokenised in parallel
// chunks (cut at line boundaries), then the SERP state machine runs over the
// events in file order, so the result is that of the sequential read.
bool parse_raw_log(const char* path, int max_sessions, std::vector<int32_t>* q_all, std::vector<int32_t>* d_all,
                   std::vector<uint8_t>* c_all, std::string* err) {
    // A regular file is mapped (the tokenising threads fault its pages in
    // parallel); anything else is read into memory.
    struct Mapping {
        void* p = MAP_FAILED;
        size_t n = 0;
        ~Mapping() { if (p != MAP_FAILED) munmap(p, n); }
    } map;
    std::string buf;
    const char* base = nullptr;
    size_t size = 0;
    {
        const int fd = open(path, O_RDONLY);
        if (fd < 0) { *err = std::string("cannot open ") + path; return false; }
        struct stat sb;
        if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0) {
            map.n = (size_t)sb.st_size;
            map.p = mmap(nullptr, map.n, PROT_READ, MAP_PRIVATE, fd, 0);
        }
        if (map.p != MAP_FAILED) {
            base = (const char*)map.p;
            size = map.n;
        } else {
            char chunk[1 << 16];
            ssize_t got;
            while ((got = read(fd, chunk, sizeof(chunk))) > 0) buf.append(chunk, (size_t)got);
            base = buf.data();
            size = buf.size();
        }
        close(fd);
    }
    const char* end = base + size;

    // Chunks of at least 4 MB, at most 16 threads.
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n_chunks = std::max<size_t>(1, std::min<size_t>(std::min<size_t>(hw, 16), size / (4u << 20)));
    std::vector<const char*> cut(n_chunks + 1, end);
    cut[0] = base;
    for (size_t c = 1; c < n_chunks; c++) {
        const char* q = base + size * c / n_chunks;
        if (q < cut[c - 1]) q = cut[c - 1];
        const char* nl = (const char*)std::memchr(q, '\n', (size_t)(end - q));
        cut[c] = nl ? nl + 1 : end;
    }
    std::vector<ChunkEvents> chunks(n_chunks);
    {
        std::vector<std::thread> th;
        for (size_t c = 1; c < n_chunks; c++) th.emplace_back(tokenize_chunk, cut[c], cut[c + 1], &chunks[c]);
        tokenize_chunk(cut[0], cut[1], &chunks[0]);
        for (auto& x : th) x.join();
    }

    {
        size_t n_query_lines = 0;
        for (const ChunkEvents& ch : chunks) n_query_lines += ch.docs.size() / kMaxSerp;
        q_all->reserve(n_query_lines);
        d_all->reserve(n_query_lines * kMaxSerp);
        c_all->reserve(n_query_lines * kMaxSerp);
    }
    int32_t cur_q = -1, cur_d[kMaxSerp];
    uint8_t cur_c[kMaxSerp];
    bool have_cur = false;
    int n_queries = 0, sid_prev = -1;
    int64_t line_base = 0;
    for (size_t c = 0; c < n_chunks; c++) {
        const ChunkEvents& ch = chunks[c];
        for (const LineEvent& e : ch.ev) {
            if (!(n_queries < max_sessions || max_sessions == -1)) return true;   // dataset.cpp:250
            if (e.kind == LineEvent::kError) {
                static const char* what[] = {"", "session id out of range", "bad query id", "bad document id",
                                             "bad clicked document id"};
                *err = std::string(what[e.value]) + " at line " + std::to_string(line_base + e.line);
                return false;
            }
            if (e.kind == LineEvent::kQuery) {
                if (sid_prev != -1 && have_cur) {                   // dataset.cpp:272-275
                    q_all->push_back(cur_q);
                    d_all->insert(d_all->end(), cur_d, cur_d + kMaxSerp);
                    c_all->insert(c_all->end(), cur_c, cur_c + kMaxSerp);
                    n_queries++;
                }
                cur_q = e.value;
                for (int r = 0; r < kMaxSerp; r++) {
                    cur_d[r] = ch.docs[(size_t)e.docs + r];
                    cur_c[r] = 0;
                }
                have_cur = true;
            } else if (e.kind == LineEvent::kClick && have_cur) {
                for (int r = 0; r < kMaxSerp; r++) {
                    if (cur_d[r] == e.value) { cur_c[r] = 1; break; }
                }
            }
            sid_prev = e.sid;                                       // dataset.cpp:288
        }
        line_base += ch.lines;
    }
    return true;
}

}  // namespace mc

// ===========================================================================
// C ABI of the dataset (include/mclicks.h)
// ===========================================================================
using namespace mc;

extern "C" {

mc_status mc_dataset_parse(const char* path, int max_sessions, float test_share, mc_dataset** out) {
    MC_REQUIRE(path && out, "null argument");
    MC_REQUIRE(max_sessions == -1 || max_sessions > 0, "max_sessions must be -1 or positive");
    std::vector<int32_t> q, d;
    std::vector<uint8_t> c;
    std::string err;
    try {
        if (!parse_raw_log(path, max_sessions, &q, &d, &c, &err)) {
            set_error(err);
            return err.rfind("cannot open", 0) == 0 ? MC_ERR_IO : MC_ERR_INVALID_ARG;
        }
        mc_dataset* ds = new mc_dataset;
        if (!prepare_dataset(ds->d, q, d, c, test_share, &err)) {
            delete ds;
            set_error(err);
            return MC_ERR_INVALID_ARG;
        }
        *out = ds;
    } catch (std::bad_alloc&) {
        set_error("out of host memory while preparing the dataset");
        return MC_ERR_OOM;
    }
    return MC_OK;
}

mc_status mc_dataset_from_sessions(int64_t n, const int32_t* query, const int32_t* docs, const uint8_t* clicks,
                                   float test_share, mc_dataset** out) {
    return mc_dataset_from_sessions_len(n, query, docs, clicks, nullptr, test_share, out);
}

mc_status mc_dataset_from_sessions_len(int64_t n, const int32_t* query, const int32_t* docs, const uint8_t* clicks,
                                       const uint8_t* len, float test_share, mc_dataset** out) {
    MC_REQUIRE(out && n > 0 && query && docs && clicks, "invalid argument");
    MC_REQUIRE(n <= INT32_MAX, "too many sessions");
    if (len) {
        for (int64_t s = 0; s < n; s++) {
            MC_REQUIRE(len[s] >= 1 && len[s] <= kMaxSerp, "len must be in [1, 10] (session " + std::to_string(s) + ")");
            for (int r = len[s]; r < kMaxSerp; r++)
                MC_REQUIRE(clicks[s * kMaxSerp + r] == 0, "click on a rank past len (session " + std::to_string(s) + ")");
        }
    }
    std::string err;
    try {
        std::vector<int32_t> q(query, query + n), d(docs, docs + n * kMaxSerp);
        std::vector<uint8_t> c(clicks, clicks + n * kMaxSerp);
        std::vector<uint8_t> l;
        if (len) l.assign(len, len + n);
        mc_dataset* ds = new mc_dataset;
        if (!prepare_dataset(ds->d, q, d, c, test_share, &err, len ? &l : nullptr)) {
            delete ds;
            set_error(err);
            return MC_ERR_INVALID_ARG;
        }
        *out = ds;
    } catch (std::bad_alloc&) {
        set_error("out of host memory while preparing the dataset");
        return MC_ERR_OOM;
    }
    return MC_OK;
}

mc_status mc_dataset_sizes(const mc_dataset* ds, int64_t* n_parsed, int64_t* n_train, int64_t* n_test,
                           int64_t* n_qd) {
    MC_REQUIRE(ds, "null dataset");
    if (n_parsed) *n_parsed = ds->d.n_parsed;
    if (n_train) *n_train = (int64_t)ds->d.train_query.size();
    if (n_test) *n_test = (int64_t)ds->d.test_query.size();
    if (n_qd) *n_qd = ds->d.n_qd;
    return MC_OK;
}

mc_status mc_dataset_export(const mc_dataset* ds, int which, int32_t* query, int32_t* docs, uint8_t* clicks,
                            int32_t* pidx) {
    MC_REQUIRE(ds && (which == 0 || which == 1), "invalid argument");
    const HostDataset& d = ds->d;
    const auto& Q = which == 0 ? d.train_query : d.test_query;
    const auto& D = which == 0 ? d.train_docs : d.test_docs;
    const auto& C = which == 0 ? d.train_clicks : d.test_clicks;
    const auto& P = which == 0 ? d.train_pidx : d.test_pidx;
    if (query) std::copy(Q.begin(), Q.end(), query);
    if (docs) std::copy(D.begin(), D.end(), docs);
    if (clicks) std::copy(C.begin(), C.end(), clicks);
    if (pidx) std::copy(P.begin(), P.end(), pidx);
    return MC_OK;
}

mc_status mc_dataset_export_len(const mc_dataset* ds, int which, uint8_t* len) {
    MC_REQUIRE(ds && (which == 0 || which == 1) && len, "invalid argument");
    const HostDataset& d = ds->d;
    const auto& L = which == 0 ? d.train_len : d.test_len;
    const size_t n = which == 0 ? d.train_query.size() : d.test_query.size();
    if (L.empty()) std::fill(len, len + n, (uint8_t)kMaxSerp);
    else std::copy(L.begin(), L.end(), len);
    return MC_OK;
}

void mc_dataset_free(mc_dataset* ds) { delete ds; }

}  // extern "C"
