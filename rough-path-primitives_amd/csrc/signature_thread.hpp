// Thread-owned subtree tiles of the batched Chen recursion
// (signature_thread.hip): one thread per (path, prefix of `split` letters).
#pragma once

#include <vector>

#include "common.hpp"

namespace rpp_hip {

template <typename T>
struct ThreadTile {
    int width;
    int depth;
    int split;    // prefix letters per thread
    int threads;  // work-group size
    int (*forward)(const DevBasis&, const rpp_hip_signature_desc&, const int64_t*, hipStream_t);
    // the backward with the same prefix length (nullptr unless split <= depth - 2)
    int (*backward)(const DevBasis&, const rpp_hip_signature_backward_desc&, const int64_t*, int64_t, hipStream_t);
};

template <typename T>
const std::vector<ThreadTile<T>>& thread_tiles();

}  // namespace rpp_hip
