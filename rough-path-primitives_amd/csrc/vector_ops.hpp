// Host-side launchers of the linear operations on batches of graded vectors
// (vector_ops.hip).
#pragma once

#include "common.hpp"

namespace rpp_hip {

enum {
    kVecAssign = 0,       // out = arg on arg's range, 0 on the rest of out's range
    kVecAdd = 1,          // out = alpha*a (on a's range) + beta*c (on c's range)
    kVecInplaceAdd = 2,   // out += alpha*a over out's range (the overlap)
    kVecScale = 3,        // out *= alpha
    kVecSetConstant = 4,  // out = alpha
    kVecSetIdentity = 5,  // out = 0, out[0] = alpha where index 0 is in range
    kVecAddIdentity = 6   // out[0] += alpha where index 0 is in range
};

// A window as the half-open range of coefficient indices [begin, end) that
// its degrees occupy; begin >= end is empty.
struct IndexRange {
    int64_t begin;
    int64_t end;
};

// One linear operation over n objects: out's coefficients in out_range are
// written, a (and c for kVecAdd) are read only inside their own ranges.
template <typename T>
int launch_vector_op(int op, int64_t n, const rpp_hip_tensor_batch& out, IndexRange out_range,
                     const rpp_hip_tensor_batch* a, IndexRange a_range, const rpp_hip_tensor_batch* c,
                     IndexRange c_range, double alpha, double beta, hipStream_t stream);

}  // namespace rpp_hip
