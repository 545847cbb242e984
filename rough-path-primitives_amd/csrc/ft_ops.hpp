// Host-side launchers of the per-operation kernels (ft_ops.hip).
#pragma once

#include "common.hpp"

namespace rpp_hip {

enum { kOpExp = 0, kOpFmexp = 1, kOpLog = 2 };

template <typename T>
int launch_product(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& lhs, const rpp_hip_tensor_batch& rhs,
                   const rpp_hip_tensor_batch* add, double alpha, double beta, hipStream_t stream);

template <typename T>
int launch_horner(int op, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                  const rpp_hip_tensor_batch& arg, const rpp_hip_tensor_batch* mult,
                  hipStream_t stream);

template <typename T>
int launch_pairing(const DevBasis& B, int64_t n, const rpp_hip_scalar_batch& out,
                   const rpp_hip_tensor_batch& f, const rpp_hip_tensor_batch& x, hipStream_t stream);

// out = L_op^* arg (left) or R_op^* arg over out's window; with accumulate the
// result is added to what out holds.
template <typename T>
int launch_adjoint(bool left, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& op, const rpp_hip_tensor_batch& arg,
                   hipStream_t stream, bool accumulate = false);

// tensor_antipode (sign_by_degree) / tensor_reflect over out's window.
template <typename T>
int launch_antipode(bool sign_by_degree, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                    const rpp_hip_tensor_batch& arg, hipStream_t stream);

// Adjoint of ft_log (log_adjoint.hip).
template <typename T>
int launch_log_adjoint(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& grad_arg,
                       const rpp_hip_tensor_batch& arg, const rpp_hip_tensor_batch& cotangent, hipStream_t stream);

// out = alpha M arg for a CSR matrix with view-relative indexing (row 0 <->
// out.begin, column c <-> arg.begin + c - col_shift), each row summed in
// entry order.  A matrix whose entries only name columns [col_shift,
// col_shift + cols) can so be applied to an argument segment that starts at
// that column; the public operation uses col_shift 0.  csc: the semantics of
// the reference's CSC product for a CSC matrix prepared as CSR
// (PreparedMatrix): alpha scales each argument value, and out rows past the
// matrix's are zeroed.
template <typename T>
int launch_spmv(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg,
                const rpp_hip_sparse_matrix& M, double alpha, hipStream_t stream, int64_t col_shift = 0,
                bool csc = false);

}  // namespace rpp_hip
