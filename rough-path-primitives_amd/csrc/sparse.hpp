// Caller-supplied sparse matrices of the sparse products (sparse.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "rpp_hip.h"

namespace rpp_hip {

// A caller's matrix in the form the CSR kernels take:
//   csr:        CSR arrays in which every row lists its entries in the order
//               the reference's CPU product adds them: the caller's own
//               arrays for a CSR matrix (entry order,
//               cpu/single_thread/operations/linalg/sparse_matrix_vector.hpp
//               :58-71), a library-owned stable transposition for a CSC
//               matrix (column order, then entry order within a column: the
//               order its column scatter reaches each row, :143-160);
//   csc:        the CSC semantics apply: alpha scales the argument before
//               each product rather than the row sum, and output rows beyond
//               the matrix's rows are zero (the whole out view is zeroed
//               first, :146).
struct PreparedMatrix {
    rpp_hip_sparse_matrix csr{};
    bool csc = false;
};

// CSR matrices pass through unchanged.  A CSC matrix is read back once (a
// host synchronisation on `stream`), checked (offsets from 0 to nnz,
// non-decreasing; row indices in [0, rows)) and transposed into device
// memory the library owns; the result is cached per (device, dtype, arrays,
// shape), so later calls issue no copies and no synchronisation.
template <typename T>
int prepare_matrix(const rpp_hip_sparse_matrix& m, PreparedMatrix* out, hipStream_t stream);

// Whether m holds exactly the library's Hall T2L (to_lie) or L2T matrix of
// (width, depth): same shape, and every row the same entries (columns and
// values in T) in column order.  Decided once per matrix and shape, from one
// read-back of the arrays.
template <typename T>
int matrix_is_hall(const rpp_hip_sparse_matrix& m, int width, int depth, bool to_lie, bool* result,
                   hipStream_t stream);

// Drops the cached preparations of every matrix whose values array is
// `values` on any device, freeing their device copies.
int forget_matrix(const void* values);

}  // namespace rpp_hip
