// mc_internal.h -- declarations shared by the host files of the EM engine.
#ifndef MC_INTERNAL_H
#define MC_INTERNAL_H

#include <cstdint>
#include <string>
#include <vector>

#include "mclicks.h"

namespace mc {

constexpr int kMaxSerp = 10;   // utils/definitions.h:50

// Prepared dataset: train set sorted by query with qd indices, filtered test set.
struct HostDataset {
    int64_t n_parsed = 0;
    int64_t n_qd = 0;
    std::vector<int32_t> train_query, train_docs, train_pidx;
    std::vector<uint8_t> train_clicks;
    std::vector<int32_t> test_query, test_docs, test_pidx;
    std::vector<uint8_t> test_clicks;
    // Ranks shown per session; empty when every session shows 10.
    std::vector<uint8_t> train_len, test_len;
};

bool parse_raw_log(const char* path, int max_sessions, std::vector<int32_t>* q_all, std::vector<int32_t>* d_all,
                   std::vector<uint8_t>* c_all, std::string* err);

bool prepare_dataset(HostDataset& ds, const std::vector<int32_t>& q_all, const std::vector<int32_t>& d_all,
                     const std::vector<uint8_t>& c_all, float test_share, std::string* err,
                     const std::vector<uint8_t>* len_all = nullptr);

void set_error(const std::string& msg);

}  // namespace mc

// Argument checks of the entry points: record the message, return a status.
#define MC_REQUIRE(cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            set_error(msg);                                                                   \
            return MC_ERR_INVALID_ARG;                                                        \
        }                                                                                     \
    } while (0)

#define MC_TRY(expr)                        \
    do {                                    \
        mc_status s_ = (expr);              \
        if (s_ != MC_OK) return s_;         \
    } while (0)

struct mc_dataset {
    mc::HostDataset d;
};

#endif  // MC_INTERNAL_H
