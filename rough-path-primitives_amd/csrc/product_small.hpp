// Compile-time-shaped products of small tensors (product_small.hip).
#pragma once

#include "common.hpp"

namespace rpp_hip {

// True when a kernel for (width, depth) is compiled (unit coefficient stride).
template <typename T>
bool product_small_available(int width, int depth);

// out = [alpha * add] + beta * lhs (x) rhs over out's window, as launch_product.
template <typename T>
int launch_product_small(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                         const rpp_hip_tensor_batch& lhs, const rpp_hip_tensor_batch& rhs,
                         const rpp_hip_tensor_batch* add, double alpha, double beta, hipStream_t stream);

}  // namespace rpp_hip
