// Two-barrier forward tiles of the batched Chen recursion (signature_chain.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace rpp_hip {

template <typename T>
struct ChainTile {
    int width;
    int depth;
    int split;    // leading letters fixed per work item
    int expand;   // levels expanded in registers per job
    int threads;  // work-group size
    int (*forward)(const DevBasis&, const rpp_hip_signature_desc&, const int64_t*, unsigned*, hipStream_t);
};

template <typename T>
const std::vector<ChainTile<T>>& chain_tiles();

}  // namespace rpp_hip
