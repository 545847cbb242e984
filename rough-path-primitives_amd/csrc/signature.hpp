// Batched Chen recursion (signature.hip): host launchers.
#pragma once

#include <vector>

#include "common.hpp"

namespace rpp_hip {

// Workspace slots of the library (see workspace()).
enum { kSlotOrder = 0, kSlotState = 1, kSlotCounter = 2, kSlotTmp = 3 };

// Reads the offsets back, checks them, and uploads the queue order (paths by
// decreasing length).  last_offset (optional) receives offsets[n], the end of
// the last path's rows: row-indexed buffers span [0, offsets[n]).
int path_order(const int64_t* d_offsets, int64_t n, std::vector<int64_t>& lengths_host, int64_t** d_order,
               hipStream_t stream, int64_t* last_offset = nullptr);

template <typename T>
int signature_forward_generic(const DevBasis& B, const rpp_hip_signature_desc& d, const int64_t* d_order,
                              hipStream_t stream);

template <typename T>
int signature_backward_generic(const DevBasis& B, const rpp_hip_signature_backward_desc& d,
                               const int64_t* d_order, hipStream_t stream);

int64_t backward_slot_elems(const DevBasis& B);

template <typename T>
int fmexp_adjoint_generic(const DevBasis& B, const rpp_hip_ft_fmexp_adjoint_desc& d, hipStream_t stream);

}  // namespace rpp_hip
