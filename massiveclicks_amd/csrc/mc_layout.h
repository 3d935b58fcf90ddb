// mc_layout.h -- host-side plan of a model's device layout (mc_layout.cpp).
#ifndef MC_LAYOUT_H
#define MC_LAYOUT_H

#include <cstdint>
#include <vector>

#include "mc_model.h"

namespace mc {

// Host arrays of a model's device layout, ready for upload.
struct HostLayout {
    std::vector<int2> counts;        // per entry: (occurrences, clicked occurrences)
    std::vector<double> sh_base;     // [2K]: shared tables' constant numerators and denominators
    std::vector<int32_t> sh_map;     // [2K]: accumulator of each, -1 = none
    std::vector<Item> items;         // work items in launch order
    std::vector<int32_t> shared_idx; // entries merged across items, ordered by minibatch
    std::vector<int32_t> lists;      // per minibatch: its window items, then its global items
    std::vector<int4> split_groups;  // per group sliced over window items: first entry, end entry, slices
    std::vector<uint32_t> packed;    // [n_pk][4] packed session records
    std::vector<uint16_t> pidx16;    // [10][n_w - n_pk] window-relative indices
    std::vector<int32_t> pidx32;     // [10][N - n_w] absolute indices
    std::vector<uint16_t> mask;      // [N] click bits
    std::vector<uint8_t> len;        // [N] ranks shown
};

// Validates the training sessions (query-grouped, session-major pidx/clicks,
// optional lengths) and plans the layout for m's model and options.  Sets
// the layout fields of m (item counts, split entries, minibatch lists,
// window widths, LDS bytes, counts, session order) and fills h.  The session
// count must already be within the stream's 32-bit offsets.
mc_status plan_layout(mc_model* m, int64_t n, const int32_t* query, const int32_t* pidx, const uint8_t* clicks,
                      const uint8_t* len, int64_t n_qd, HostLayout* h);

}  // namespace mc

#endif  // MC_LAYOUT_H
