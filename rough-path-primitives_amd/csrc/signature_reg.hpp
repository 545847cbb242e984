// Register-resident tiles of the batched Chen recursion (signature_reg.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace rpp_hip {

template <typename T>
struct RegTile {
    int dtype_bytes;
    int width;
    int depth;
    int threads;  // work-group size
    int split;    // leading letters fixed per work-group
    int reg_levels;  // top subtree levels held in registers (the rest in LDS)
    int (*forward)(const DevBasis&, const rpp_hip_signature_desc&, const int64_t*, unsigned*, hipStream_t);
    int (*backward)(const DevBasis&, const rpp_hip_signature_backward_desc&, const int64_t*, int64_t, unsigned*,
                    hipStream_t);
    // per-operation adjoint of that call (split <= 1)
    int (*fmexp_adjoint)(const DevBasis&, const rpp_hip_ft_fmexp_adjoint_desc&, hipStream_t);
};

// Every compiled tile for dtype T; the first match of a (width, depth) is the
// default when a descriptor names no tile.
template <typename T>
const std::vector<RegTile<T>>& reg_tiles();

template <typename T>
const RegTile<T>* fmexp_adj_tile(int width, int depth);

}  // namespace rpp_hip
