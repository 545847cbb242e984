// ft_log with the object resident in LDS (log_lds.hip).
#pragma once

#include "common.hpp"

namespace rpp_hip {

// True when a kernel for (width, depth) is compiled and one object's argument
// and result fit in LDS for this dtype.
template <typename T>
bool log_lds_available(int width, int depth);

// out = log(1 + arg_{>=1}) with out's window [0, depth], unit coefficient
// strides, out not aliasing arg.
template <typename T>
int launch_log_lds(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg,
                   hipStream_t stream);

}  // namespace rpp_hip
