// Streaming tiles of the plain ft_fmexp call (fmexp_coal.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace rpp_hip {

template <typename T>
struct FmexpTile {
    int kind;  // 2: coalesced subtree items (k_fmexp_coal), 3: sibling-group wavefronts (k_fmexp_wave)
    int dtype_bytes;
    int width;
    int depth;
    int threads;        // work-group size
    int split;          // leading letters fixed per work item (kind 2)
    int blocks_per_cu;  // grid cap per compute unit (0: one work-group per item)
    int (*launch)(const DevBasis&, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& mult,
                  const rpp_hip_tensor_batch& ex, int64_t n, hipStream_t);
};

template <typename T>
const std::vector<FmexpTile<T>>& fmexp_coal_tiles();

// The tile serving plain ft_fmexp at (width, depth), or nullptr.
template <typename T>
const FmexpTile<T>* fmexp_coal_tile(int width, int depth);

}  // namespace rpp_hip
