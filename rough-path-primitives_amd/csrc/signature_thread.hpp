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
    // the backward with the same prefix length (nullptr unless split <= depth - 2);
    // partial: device buffer of at least thread_backward_partial_bytes(), or
    // nullptr for the library workspace
    int (*backward)(const DevBasis&, const rpp_hip_signature_backward_desc&, const int64_t*, int64_t, hipStream_t,
                    void* partial);
};

template <typename T>
const std::vector<ThreadTile<T>>& thread_tiles();

// Bytes of the backward's per-row partial sums: rows [0, total_rows) of
// width * width values.
template <typename T>
size_t thread_backward_partial_bytes(int width, int64_t total_rows) {
    return static_cast<size_t>(total_rows > 0 ? total_rows : 1) * static_cast<size_t>(width) *
           static_cast<size_t>(width) * sizeof(T);
}

// The fixed device scratch of the allocation-free backward
// (rpp_hip_signature_backward_noalloc_*): a file-scope __device__ array of
// kNoAllocScratchBytes; *ptr receives its address on the current device
// (hipGetSymbolAddress once per device, then cached).
constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;
int noalloc_scratch(void** ptr);

}  // namespace rpp_hip
