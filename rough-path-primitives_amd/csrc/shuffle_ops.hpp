// Host-side launcher of the shuffle-family kernels (shuffle_ops.hip).
#pragma once

#include "common.hpp"

namespace rpp_hip {

enum {
    kStMul = 0,      // out = beta b.c                      (p = b, q = c)
    kStFma = 1,      // out = alpha a + beta b.c            (r = a, p = b, q = c)
    kStInplace = 2,  // out = alpha out + beta b.c          (p = b, q = c)
    kStAdjMul = 3,   // <out, x> = <q, p.x>                 (p = op, q = arg)
    kHsMul = 4,      // out = beta p<q                      (p = integrator, q = integrand)
    kHsAdjL = 5,     // <out, z> = beta <q, p<z>
    kHsAdjR = 6      // <out, z> = beta <p, z<q>
};

// Words longer than this are refused: a shuffle visits 2^degree letter
// subsets per output coefficient.
constexpr int kShuffleMaxDepth = 16;

template <typename T>
int launch_shuffle(int op, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& p, const rpp_hip_tensor_batch& q, const rpp_hip_tensor_batch* r,
                   double alpha, double beta, hipStream_t stream);

}  // namespace rpp_hip
