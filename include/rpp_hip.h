/*
 * rpp_hip.h -- C-ABI of librpp_hip.so, the MI355X (gfx950) strategy for the
 * truncated tensor algebra hot path of RoughPathPrimitives.
 *
 * Every entry point is asynchronous on `stream` (a hipStream_t passed as a
 * plain pointer; NULL is the default stream), exactly like the reference GPU
 * strategy, whose launch() enqueues on LaunchConfig.stream and returns
 * (gpu/device.hpp:21-24, gpu/block/strategy.hpp:312-365).  All tensor data are
 * device pointers owned by the caller; the library owns only its per-device
 * caches (Lie-basis matrices per (width, depth), workspaces), which are
 * created on first use and kept until unload.
 *
 * Return values are the integer codes of rpp::ErrorCode
 * (support/error.hpp:28-39).  On a non-zero return, rpp_hip_last_error()
 * gives a message for the calling thread.  Descriptors are validated on the
 * host before any HIP call.
 */
#ifndef RPP_HIP_H
#define RPP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same numbering as rpp::ErrorCode (support/error.hpp:28-39). */
enum rpp_hip_status {
    RPP_HIP_SUCCESS = 0,
    RPP_HIP_CANCELLED = 1,
    RPP_HIP_UNKNOWN = 2,
    RPP_HIP_INVALID_ARGUMENT = 3,
    RPP_HIP_TIMEOUT = 4,
    RPP_HIP_OUT_OF_RESOURCES = 5,
    RPP_HIP_CONTRACT_VIOLATION = 6,
    RPP_HIP_OUT_OF_BOUNDS = 7,
    RPP_HIP_NOT_IMPLEMENTED = 8,
    RPP_HIP_INTERNAL = 9
};

/* TensorBasis(width, depth) (basis/tensor_basis.hpp:40-118).  Degree offsets
 * start_of_degree(k) = (width^k - 1)/(width - 1) are implied. */
typedef struct rpp_hip_basis {
    int32_t width;
    int32_t depth;
} rpp_hip_basis;

/* One Batch of DenseTensorView (views/batch.hpp:101-145,
 * views/dense_tensor_view.hpp:38-91): object b, coefficient i lives at
 *   data[b * batch_stride + i * coeff_stride].
 * coeff_stride == 1 is layouts::StrideLayout (object-major); batch_stride == 1
 * with coeff_stride == num_batches is the batch-interleaved layout; a
 * batch_stride of 0 broadcasts one object to the whole batch.  Only degrees
 * [min_degree, max_degree] are read or written (the view's window). */
typedef struct rpp_hip_tensor_batch {
    void* data;
    int64_t batch_stride;
    int64_t coeff_stride;
    int32_t min_degree;
    int32_t max_degree;
} rpp_hip_tensor_batch;

/* Aliasing.  Two views "share memory" when some element the one touches
 * (its window, over all num_batches objects) is an element the other
 * touches.  Where an entry point below lets an output be an input, it must be
 * the same batch (same data, batch_stride and coeff_stride); every other
 * shared element is rejected with RPP_HIP_INVALID_ARGUMENT before any launch.
 * The test is exact for two object-major or two batch-interleaved views with
 * equal strides (for example disjoint column slices of one buffer), and
 * conservative otherwise: views with different strides whose byte ranges
 * intersect count as sharing memory. */

/* A batch of ScalarView (views/scalar_view.hpp:17-60): object b at
 * data[b * stride]; make_scalar_batch uses stride 1. */
typedef struct rpp_hip_scalar_batch {
    void* data;
    int64_t stride;
} rpp_hip_scalar_batch;

/* ---------------------------------------------------------------- products */

/* rpp::ops::ft_mul (operations/basic/ft_mul.hpp:53-79):
 *   out = beta * lhs (x) rhs on out's window; out may be lhs or rhs (the same
 *   batch) and shares memory with neither otherwise. */
typedef struct rpp_hip_ft_mul_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, lhs, rhs;
    double beta;
} rpp_hip_ft_mul_desc;
int rpp_hip_ft_mul_f32(const rpp_hip_ft_mul_desc* desc, void* stream);
int rpp_hip_ft_mul_f64(const rpp_hip_ft_mul_desc* desc, void* stream);

/* rpp::ops::ft_fma (operations/basic/ft_fma.hpp:54-83):
 *   out = alpha * a + beta * b (x) c; a is read only inside its window. */
typedef struct rpp_hip_ft_fma_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, a, b, c;
    double alpha, beta;
} rpp_hip_ft_fma_desc;
int rpp_hip_ft_fma_f32(const rpp_hip_ft_fma_desc* desc, void* stream);
int rpp_hip_ft_fma_f64(const rpp_hip_ft_fma_desc* desc, void* stream);

/* rpp::ops::ft_inplace_fma<FMAType> (operations/basic/ft_inplace_fma.hpp
 * :15-19,64-100).  variant 231: a <- alpha*a + beta*b(x)c,
 * 123: a <- alpha*c + beta*a(x)b, 213: a <- alpha*c + beta*b(x)a. */
typedef struct rpp_hip_ft_inplace_fma_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    int32_t variant;
    rpp_hip_tensor_batch a, b, c;
    double alpha, beta;
} rpp_hip_ft_inplace_fma_desc;
int rpp_hip_ft_inplace_fma_f32(const rpp_hip_ft_inplace_fma_desc* desc, void* stream);
int rpp_hip_ft_inplace_fma_f64(const rpp_hip_ft_inplace_fma_desc* desc, void* stream);

/* rpp::ops::ft_inplace_mul (operations/basic/ft_inplace_mul.hpp:38-69):
 *   lhs <- beta * lhs (x) rhs. */
typedef struct rpp_hip_ft_inplace_mul_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch lhs, rhs;
    double beta;
} rpp_hip_ft_inplace_mul_desc;
int rpp_hip_ft_inplace_mul_f32(const rpp_hip_ft_inplace_mul_desc* desc, void* stream);
int rpp_hip_ft_inplace_mul_f64(const rpp_hip_ft_inplace_mul_desc* desc, void* stream);

/* ------------------------------------------------------------- composites */

/* rpp::ops::ft_exp (operations/intermediate/ft_exp.hpp:89-112):
 *   out = exp(arg_{>=1}); and rpp::ops::ft_log (ft_log.hpp:91-114):
 *   out = log(1 + arg_{>=1}).  One kernel per call. */
typedef struct rpp_hip_ft_unary_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, arg;
} rpp_hip_ft_unary_desc;
int rpp_hip_ft_exp_f32(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_ft_exp_f64(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_ft_log_f32(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_ft_log_f64(const rpp_hip_ft_unary_desc* desc, void* stream);

/* rpp::ops::ft_fmexp (operations/intermediate/ft_fmexp.hpp:96-121):
 *   out = multiplier (x) exp(exponent_{>=1}); out shares memory with neither input.
 *   One kernel per call. */
typedef struct rpp_hip_ft_fmexp_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, multiplier, exponent;
} rpp_hip_ft_fmexp_desc;
int rpp_hip_ft_fmexp_f32(const rpp_hip_ft_fmexp_desc* desc, void* stream);
int rpp_hip_ft_fmexp_f64(const rpp_hip_ft_fmexp_desc* desc, void* stream);

/* rpp::ops::tensor_pairing (operations/basic/tensor_pairing.hpp:47-73):
 *   out[b] = sum over the overlap of the windows of functional[i]*arg[i]. */
typedef struct rpp_hip_tensor_pairing_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_scalar_batch out;
    rpp_hip_tensor_batch functional, arg;
} rpp_hip_tensor_pairing_desc;
int rpp_hip_tensor_pairing_f32(const rpp_hip_tensor_pairing_desc* desc, void* stream);
int rpp_hip_tensor_pairing_f64(const rpp_hip_tensor_pairing_desc* desc, void* stream);

/* rpp::ops::ft_adj_lmul (operations/basic/ft_adj_lmul.hpp:49-73):
 *   out[w] = sum_u op[u] arg[uw];  rpp::ops::ft_adj_rmul (ft_adj_rmul.hpp
 *   :47-71): out[w] = sum_u arg[wu] op[u]. */
typedef struct rpp_hip_ft_adjoint_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, op, arg;
} rpp_hip_ft_adjoint_desc;
int rpp_hip_ft_adj_lmul_f32(const rpp_hip_ft_adjoint_desc* desc, void* stream);
int rpp_hip_ft_adj_lmul_f64(const rpp_hip_ft_adjoint_desc* desc, void* stream);
int rpp_hip_ft_adj_rmul_f32(const rpp_hip_ft_adjoint_desc* desc, void* stream);
int rpp_hip_ft_adj_rmul_f64(const rpp_hip_ft_adjoint_desc* desc, void* stream);

/* Adjoint of rpp::ops::ft_fmexp for a degree-one exponent (one Chen step
 * S' = S (x) exp(x)): given the multiplier S, the exponent x (degree 1 of a
 * tensor batch) and the cotangent G = d loss / d out, write
 *   grad_multiplier = R_{exp(x)}^* G   (the transpose of right multiplication,
 *                                       as rpp::ops::ft_adj_rmul with
 *                                       op = exp(x), ft_adj_rmul.hpp:47-71)
 *   grad_exponent   = the x-gradient, into degree 1 of a tensor batch.
 * Full windows [0, depth] for multiplier, cotangent and grad_multiplier; the
 * exponent and grad_exponent windows must hold degree 1.  No reference
 * counterpart (the reference has no backward for ft_fmexp). */
typedef struct rpp_hip_ft_fmexp_adjoint_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch multiplier, exponent, cotangent;
    rpp_hip_tensor_batch grad_multiplier, grad_exponent;
} rpp_hip_ft_fmexp_adjoint_desc;
int rpp_hip_ft_fmexp_adjoint_f32(const rpp_hip_ft_fmexp_adjoint_desc* desc, void* stream);
int rpp_hip_ft_fmexp_adjoint_f64(const rpp_hip_ft_fmexp_adjoint_desc* desc, void* stream);

/* Adjoint of rpp::ops::ft_log (operations/intermediate/ft_log.hpp:66-85):
 * given arg and the cotangent G = d loss / d log(arg) (log taken over the
 * full window [0, depth]; G's window is zero-extended), write
 * grad_arg = d loss / d arg on arg's window (degree 0 gets 0: ft_log reads
 * arg from degree 1).  grad_arg's window must equal arg's and it must not
 * share memory with arg or the cotangent.  Uses a library workspace of about
 * (3 + 1/(width-1)) * dim coefficients per object.  No reference counterpart
 * (the reference has no backward for ft_log). */
typedef struct rpp_hip_ft_log_adjoint_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch arg, cotangent, grad_arg;
} rpp_hip_ft_log_adjoint_desc;
int rpp_hip_ft_log_adjoint_f32(const rpp_hip_ft_log_adjoint_desc* desc, void* stream);
int rpp_hip_ft_log_adjoint_f64(const rpp_hip_ft_log_adjoint_desc* desc, void* stream);

/* rpp::ops::tensor_antipode and rpp::ops::tensor_reflect
 * (operations/basic/tensor_antipode.hpp:15-39, tensor_reflect.hpp:15-39;
 * TensorGeneralisedAntipode, tensor_generalised_antipode.hpp:17-81): on each
 * degree k of out's window, out[reverse(w)] = (-1)^k arg[w] (antipode) or
 * arg[w] (reflect); 0 where arg's window lacks degree k.  out must not share
 * memory with arg. */
int rpp_hip_tensor_antipode_f32(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_tensor_antipode_f64(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_tensor_reflect_f32(const rpp_hip_ft_unary_desc* desc, void* stream);
int rpp_hip_tensor_reflect_f64(const rpp_hip_ft_unary_desc* desc, void* stream);

/* ------------------------------------------------------ linear operations */

/* The graded basis of a vector operation: the TensorBasis of (width, depth)
 * (basis/tensor_basis.hpp:40-130) or the Hall-set LieBasis of (width, depth)
 * (basis/lie_basis.hpp:42-97: degree 0 holds one slot, degree k the Witt
 * number of Lie elements).  These operations use only its degree offsets. */
enum rpp_hip_basis_kind { RPP_HIP_TENSOR_BASIS = 0, RPP_HIP_LIE_BASIS = 1 };
typedef struct rpp_hip_graded_basis {
    int32_t kind;
    int32_t width;
    int32_t depth;
} rpp_hip_graded_basis;

/* rpp::ops::vector_assign (operations/linalg/vector_assign.hpp:36-64;
 * cpu/.../linalg/vector_assign.hpp:32-61): out = arg on the overlap of the two
 * windows, 0 on the rest of out's window.  out and arg may have different
 * layouts: between an object-major and a batch-interleaved batch this is the
 * transposition of the two wire formats (an LDS-tiled kernel).  out may be arg
 * itself (same data and strides); other overlaps are not supported. */
typedef struct rpp_hip_vector_assign_desc {
    rpp_hip_graded_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, arg;
} rpp_hip_vector_assign_desc;
int rpp_hip_vector_assign_f32(const rpp_hip_vector_assign_desc* desc, void* stream);
int rpp_hip_vector_assign_f64(const rpp_hip_vector_assign_desc* desc, void* stream);

/* rpp::ops::vector_add (operations/linalg/vector_add.hpp:40-78;
 * cpu/.../linalg/vector_add.hpp:36-66): on every degree of out's window,
 * out = alpha*lhs + beta*rhs, each operand read only on its own window and 0
 * elsewhere.  out may be lhs or rhs (same data and strides). */
typedef struct rpp_hip_vector_add_desc {
    rpp_hip_graded_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, lhs, rhs;
    double alpha, beta;
} rpp_hip_vector_add_desc;
int rpp_hip_vector_add_f32(const rpp_hip_vector_add_desc* desc, void* stream);
int rpp_hip_vector_add_f64(const rpp_hip_vector_add_desc* desc, void* stream);

/* rpp::ops::vector_inplace_add (operations/linalg/vector_inplace_add.hpp
 * :40-77; cpu/.../linalg/vector_inplace_add.hpp:36-63): lhs += alpha*rhs on
 * the overlap of the two windows; the rest of lhs is untouched. */
typedef struct rpp_hip_vector_inplace_add_desc {
    rpp_hip_graded_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch lhs, rhs;
    double alpha;
} rpp_hip_vector_inplace_add_desc;
int rpp_hip_vector_inplace_add_f32(const rpp_hip_vector_inplace_add_desc* desc, void* stream);
int rpp_hip_vector_inplace_add_f64(const rpp_hip_vector_inplace_add_desc* desc, void* stream);

/* One batch and one scalar, over the batch's window:
 *   vector_scalar_multiply (operations/linalg/vector_scalar_multiply.hpp
 *     :28-71; cpu/.../vector_scalar_multiply.hpp:37-49): vec *= scalar;
 *   vector_set_constant (operations/linalg/vector_set_constant.hpp:27-72;
 *     cpu/.../vector_set_constant.hpp:34-40): vec = scalar;
 *   tensor_set_identity (operations/basic/tensor_set_identity.hpp:27-73;
 *     cpu/.../tensor_set_identity.hpp:34-44): vec = 0, then vec[0] = scalar
 *     if the window holds degree 0;
 *   tensor_add_identity (operations/basic/tensor_add_identity.hpp:26-71;
 *     cpu/.../tensor_add_identity.hpp:34-43): vec[0] += scalar if the window
 *     holds degree 0. */
typedef struct rpp_hip_vector_scalar_desc {
    rpp_hip_graded_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch vec;
    double scalar;
} rpp_hip_vector_scalar_desc;
int rpp_hip_vector_scalar_multiply_f32(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_vector_scalar_multiply_f64(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_vector_set_constant_f32(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_vector_set_constant_f64(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_tensor_set_identity_f32(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_tensor_set_identity_f64(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_tensor_add_identity_f32(const rpp_hip_vector_scalar_desc* desc, void* stream);
int rpp_hip_tensor_add_identity_f64(const rpp_hip_vector_scalar_desc* desc, void* stream);

/* ---------------------------------------------------- the shuffle family */

/* Shuffle products on the dual (functional) side.  For a word w of degree k,
 * (b . c)[w] = sum over the subsets S of w's positions of b[w|S] c[w|S^c],
 * each factor read only on its window; the left half-shuffle (x < y)[a w] =
 * sum over S of x[a w|S] y[w|S^c] fixes the first letter on x's side.
 * Depth at most 16 (2^k subsets per coefficient); outputs must not share
 * memory with inputs.
 *   rpp_hip_st_mul: rpp::ops::st_mul (operations/basic/st_mul.hpp:48-83,
 *     cpu/.../basic/st_mul.hpp:37-66): out = beta lhs . rhs.
 *   rpp_hip_st_fma: rpp::ops::st_fma (operations/basic/st_fma.hpp:50-89,
 *     cpu/.../st_fma.hpp:41-79): out = alpha a + beta b . c, a read only on
 *     its window; out may be a itself (same data and strides). */
int rpp_hip_st_mul_f32(const rpp_hip_ft_mul_desc* desc, void* stream);
int rpp_hip_st_mul_f64(const rpp_hip_ft_mul_desc* desc, void* stream);
int rpp_hip_st_fma_f32(const rpp_hip_ft_fma_desc* desc, void* stream);
int rpp_hip_st_fma_f64(const rpp_hip_ft_fma_desc* desc, void* stream);

/* rpp::ops::st_inplace_fma (operations/basic/st_inplace_fma.hpp:45-85,
 * cpu/.../st_inplace_fma.hpp:38-68): a <- alpha a + beta b . c on a's
 * window; b and c must not share memory with a. */
typedef struct rpp_hip_st_inplace_fma_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch a, b, c;
    double alpha, beta;
} rpp_hip_st_inplace_fma_desc;
int rpp_hip_st_inplace_fma_f32(const rpp_hip_st_inplace_fma_desc* desc, void* stream);
int rpp_hip_st_inplace_fma_f64(const rpp_hip_st_inplace_fma_desc* desc, void* stream);

/* rpp::ops::st_adj_mul (operations/basic/st_adj_mul.hpp:40-73,
 * cpu/.../st_adj_mul.hpp:43-82): the adjoint of shuffle multiplication by
 * op, <out, x> = <arg, op . x>, on out's window. */
int rpp_hip_st_adj_mul_f32(const rpp_hip_ft_adjoint_desc* desc, void* stream);
int rpp_hip_st_adj_mul_f64(const rpp_hip_ft_adjoint_desc* desc, void* stream);

/* The left half-shuffle and its adjoints:
 *   rpp_hip_left_hs_mul: rpp::ops::left_hs_mul (operations/basic/
 *     left_hs_mul.hpp:53-93, cpu/.../left_hs_mul.hpp:37-84):
 *     out = beta integrator < integrand; degree 0 of out is 0.
 *   rpp_hip_left_hs_adj_lmul: rpp::ops::left_hs_adj_lmul (operations/basic/
 *     left_hs_adj_lmul.hpp:33-61, gpu/block/operations/basic/
 *     left_hs_adj_lmul.hpp:114-177): <out, z> = beta <integrand, integrator < z>.
 *   rpp_hip_left_hs_adj_rmul: rpp::ops::left_hs_adj_rmul (left_hs_adj_rmul.hpp
 *     :33-61, gpu/.../left_hs_adj_rmul.hpp:77-143): <out, z> =
 *     beta <integrator, z < integrand>; degree 0 of out is 0.
 * beta scales the adjoints as the reference's GPU operators do (its CPU
 * operators compute beta = 1). */
typedef struct rpp_hip_half_shuffle_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch out, integrator, integrand;
    double beta;
} rpp_hip_half_shuffle_desc;
int rpp_hip_left_hs_mul_f32(const rpp_hip_half_shuffle_desc* desc, void* stream);
int rpp_hip_left_hs_mul_f64(const rpp_hip_half_shuffle_desc* desc, void* stream);
int rpp_hip_left_hs_adj_lmul_f32(const rpp_hip_half_shuffle_desc* desc, void* stream);
int rpp_hip_left_hs_adj_lmul_f64(const rpp_hip_half_shuffle_desc* desc, void* stream);
int rpp_hip_left_hs_adj_rmul_f32(const rpp_hip_half_shuffle_desc* desc, void* stream);
int rpp_hip_left_hs_adj_rmul_f64(const rpp_hip_half_shuffle_desc* desc, void* stream);

/* --------------------------------------------------------- sparse and Lie */

enum rpp_hip_matrix_format { RPP_HIP_CSR = 0, RPP_HIP_CSC = 1 };

/* CompressedMatrix (sparse/detail/compressed_matrix.hpp:38-126) with device
 * arrays: values[nnz], indices[nnz] (inner index), offsets[outer + 1]. */
typedef struct rpp_hip_sparse_matrix {
    int32_t format;
    const void* values;      /* float or double, as the entry point */
    const int64_t* indices;
    const int64_t* offsets;
    int64_t nnz;
    int64_t rows;
    int64_t cols;
} rpp_hip_sparse_matrix;

/* A batch of vectors for SpMV: object b, entry i of the view at
 * data[b*batch_stride + (begin + i)*coeff_stride], where begin is
 * view.begin_index() and size the view's length (R9: row 0 <-> out.begin(),
 * column 0 <-> arg.begin()). */
typedef struct rpp_hip_vector_batch {
    void* data;
    int64_t batch_stride;
    int64_t coeff_stride;
    int64_t begin;
    int64_t size;
} rpp_hip_vector_batch;

/* rpp::ops::sparse_matrix_vector_product
 * (operations/linalg/sparse_matrix_vector.hpp:57-96): out = alpha * M arg,
 * every row summed in the order the reference's cpu::single_thread product
 * sums it, with no atomics, so results are deterministic and bit-identical to
 * the reference:
 *   CSR: entry order, out = alpha * sum (cpu/.../linalg/
 *        sparse_matrix_vector.hpp:45-73);
 *   CSC: the out view is zeroed and each row receives value * (alpha * arg)
 *        in column order (:133-163).  A CSC matrix is read back once (one
 *        host synchronisation on the stream), checked (offsets 0..nnz,
 *        non-decreasing; row indices in range) and transposed stably into
 *        device memory the library owns; that form is cached per (device,
 *        dtype, the three array pointers, shape), so later calls issue kernel
 *        launches only.  The library assumes the arrays of a cached matrix do
 *        not change; after rewriting or freeing them call
 *        rpp_hip_sparse_matrix_forget(values). */
typedef struct rpp_hip_spmv_desc {
    int64_t num_batches;
    rpp_hip_vector_batch out, arg;
    rpp_hip_sparse_matrix matrix;
    double alpha;
} rpp_hip_spmv_desc;
int rpp_hip_sparse_matrix_vector_product_f32(const rpp_hip_spmv_desc* desc, void* stream);
int rpp_hip_sparse_matrix_vector_product_f64(const rpp_hip_spmv_desc* desc, void* stream);

/* Drops what the library cached for every matrix whose values array is
 * `values` (the CSR form of a CSC matrix, the Hall-matrix recognition of the
 * *_matrix Lie conversions below), synchronising the devices that hold such
 * a form before freeing it.  No reference counterpart: the reference
 * re-reads its matrices on every call. */
int rpp_hip_sparse_matrix_forget(const void* values);

/* rpp::ops::tensor_to_lie (operations/basic/tensor_to_lie.hpp:65-103) and
 * rpp::ops::lie_to_tensor (operations/basic/lie_to_tensor.hpp:70-108) with the
 * library's own Hall-basis T2L / L2T matrices, built on the host and
 * uploaded once per (width, depth, dtype) as CSR.  Lie objects hold
 * true_size() slots (slot 0 unused) and Lie windows start at degree >= 1. */
typedef struct rpp_hip_lie_convert_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch tensor;
    rpp_hip_tensor_batch lie;
} rpp_hip_lie_convert_desc;
int rpp_hip_tensor_to_lie_f32(const rpp_hip_lie_convert_desc* desc, void* stream);
int rpp_hip_tensor_to_lie_f64(const rpp_hip_lie_convert_desc* desc, void* stream);
int rpp_hip_lie_to_tensor_f32(const rpp_hip_lie_convert_desc* desc, void* stream);
int rpp_hip_lie_to_tensor_f64(const rpp_hip_lie_convert_desc* desc, void* stream);

/* rpp::ops::tensor_to_lie (operations/basic/tensor_to_lie.hpp:71-78) and
 * rpp::ops::lie_to_tensor (operations/basic/lie_to_tensor.hpp:76-83) with the
 * caller's matrix, as the reference's free functions take it (typically the
 * CSC output of make_t2l_matrix / make_l2t_matrix, sparse/t2l_matrix.hpp:57,
 * sparse/l2t_matrix.hpp:58): out = M arg relative to the view begins.  When
 * the windows are the reference's (tensor [0, depth], Lie [1, depth]) and M
 * equals the library's Hall T2L (L2T) of the basis (decided once per matrix
 * from one read-back, then cached), the library's Lie kernel serves the call;
 * otherwise the sparse product above (the matrix must fit the views).  Both
 * sum each row in the reference's order, so the result is the same bits. */
typedef struct rpp_hip_lie_matrix_convert_desc {
    rpp_hip_basis basis;
    int64_t num_batches;
    rpp_hip_tensor_batch tensor;
    rpp_hip_tensor_batch lie;
    rpp_hip_sparse_matrix matrix;
} rpp_hip_lie_matrix_convert_desc;
int rpp_hip_tensor_to_lie_matrix_f32(const rpp_hip_lie_matrix_convert_desc* desc, void* stream);
int rpp_hip_tensor_to_lie_matrix_f64(const rpp_hip_lie_matrix_convert_desc* desc, void* stream);
int rpp_hip_lie_to_tensor_matrix_f32(const rpp_hip_lie_matrix_convert_desc* desc, void* stream);
int rpp_hip_lie_to_tensor_matrix_f64(const rpp_hip_lie_matrix_convert_desc* desc, void* stream);

/* Host-side Hall basis metadata (basis/hall_basis.hpp:16-88,
 * sparse/l2t_matrix.hpp:58-114, sparse/t2l_matrix.hpp:57-142).  Sizes first:
 * sizes[0] = true_size (incl. slot 0), [1] = L2T nnz, [2] = T2L nnz,
 * [3] = tensor dimension.  Then the arrays: pairs[2*true_size],
 * degree_begin[depth+2], and both matrices in CSC exactly as the reference
 * builders return them (L2T: columns = Lie keys 1.., rows = tensor words;
 * T2L: columns = tensor words, rows = Lie key - 1).  Any output pointer may be
 * NULL. */
int rpp_hip_lie_basis_sizes(int32_t width, int32_t depth, int64_t* sizes);
int rpp_hip_lie_basis_copy(int32_t width, int32_t depth, int64_t* pairs,
                           int64_t* degree_begin, double* l2t_values,
                           int64_t* l2t_indices, int64_t* l2t_offsets,
                           double* t2l_values, int64_t* t2l_indices,
                           int64_t* t2l_offsets);

/* The Hall basis metadata on disk.  No counterpart in the reference, whose
 * builders (basis/hall_basis.hpp, sparse/l2t_matrix.hpp, sparse/t2l_matrix.hpp)
 * run again in every process; for callers that build large shapes once and
 * reuse them across processes.
 *   rpp_hip_lie_basis_save: writes the metadata of (width, depth), building it
 *     if this process has not, to `path` (a little-endian binary file: format
 *     tag, shape, sizes, the arrays of rpp_hip_lie_basis_copy in double and
 *     int64, an FNV-1a checksum), atomically (temporary file, then rename).
 *   rpp_hip_lie_basis_load: reads such a file, checks the checksum, every
 *     size and every index range, and makes it this process's metadata for its
 *     (width, depth), returned through width/depth (either may be NULL).  A
 *     file that cannot be opened is RPP_HIP_INVALID_ARGUMENT; a damaged or
 *     inconsistent one RPP_HIP_CONTRACT_VIOLATION, as is a file that differs
 *     from the metadata this process already uses for that shape.
 * With the environment variable RPP_HIP_LIE_CACHE_DIR set to a directory, the
 * library does the same by itself: the first use of a shape reads
 * rpp_lie_w<width>_d<depth>.bin there, or builds the metadata and writes that
 * file. */
int rpp_hip_lie_basis_save(int32_t width, int32_t depth, const char* path);
int rpp_hip_lie_basis_load(const char* path, int32_t* width, int32_t* depth);

/* ------------------------------------------------- batched Chen recursion */

/* Signatures of many piecewise-linear paths of uneven length in ONE
 * persistent launch (the caller-side chain of ft_fmexp of the reference's
 * usage pattern: S_0 = 1, S_t = S_{t-1} (x) exp(dX_t)).
 *   increments: device [total, width] row-major; path p owns rows
 *               offsets[p] .. offsets[p+1]-1 (offsets: device int64[n+1]).
 *   out:        tensor batch with the full window [0, depth].
 *   checkpoint_every / checkpoints: optional (0 / NULL).  When set, the
 *               state after every checkpoint_every-th increment of path p is
 *               stored at checkpoints + (ckpt_offsets[p] + j) * dim, used by
 *               the backward pass to re-anchor its reconstruction.  Degrees
 *               0 .. depth-1 are stored; the top degree, which no backward
 *               step reads, is left unwritten by the thread-owned subtree
 *               tiles (tile kind 2).
 *   order / num_increments: optional.  NULL order: the library reads the
 *               offsets back (one host synchronisation) and queues paths by
 *               decreasing length.  Otherwise order is a device permutation
 *               of 0..n-1 (the queue order, longest paths first for balance)
 *               and num_increments >= offsets[n] (the row count of
 *               increments, for instance): the call then issues kernel
 *               launches only, so it can be captured in a hipGraph once the
 *               library workspace has been sized by an uncaptured call.
 * tile_threads > 0 selects a compiled tile (rpp_hip_signature_tiles),
 * 0 the library default for (width, depth), < 0 the storage-generic kernel. */
typedef struct rpp_hip_signature_desc {
    rpp_hip_basis basis;
    int64_t num_paths;
    const void* increments;
    const int64_t* offsets;
    rpp_hip_tensor_batch out;
    int32_t checkpoint_every;
    void* checkpoints;
    const int64_t* ckpt_offsets;
    int32_t tile_threads;
    int32_t tile_split;
    int32_t tile_kind;    /* 0: Horner register tile, 1: two-barrier chain tile,
                             2: thread-owned subtree tile */
    int32_t tile_expand;  /* chain tiles: levels expanded per job; register
                             tiles: top levels held in registers (0: any) */
    const int64_t* order;
    int64_t num_increments;
} rpp_hip_signature_desc;
int rpp_hip_signature_f32(const rpp_hip_signature_desc* desc, void* stream);
int rpp_hip_signature_f64(const rpp_hip_signature_desc* desc, void* stream);

/* Backward of rpp_hip_signature_*: given the final signatures S_L (out of the
 * forward pass) and the cotangent G_L = d loss / d S_L (a tensor batch, or one
 * functional broadcast with batch_stride 0), write d loss / d increments
 * (device [total, width]): rows offsets[0] .. offsets[n]-1 are written, rows
 * outside that range (which belong to no path) are left untouched; offsets[0]
 * may be above 0.  Intermediate states are reconstructed backward
 * with S_{t-1} = S_t (x) exp(-dX_t); when the forward pass stored checkpoints
 * the reconstruction is re-anchored at each of them (sqrt(L) checkpointing).
 * scratch: device workspace of rpp_hip_signature_backward_scratch_bytes() for
 * the storage-generic kernel (NULL: library workspace).  tile_threads /
 * tile_split select a register tile as for the forward pass (split <= 1), or
 * with tile_kind 2 a thread-owned subtree tile that has a backward (its
 * work-group is one (path, first letter) with one thread per prefix of
 * tile_split letters); order / num_increments as for the forward pass. */
typedef struct rpp_hip_signature_backward_desc {
    rpp_hip_basis basis;
    int64_t num_paths;
    const void* increments;
    const int64_t* offsets;
    rpp_hip_tensor_batch signatures;
    rpp_hip_tensor_batch cotangent;
    void* grad_increments;
    int32_t checkpoint_every;
    const void* checkpoints;
    const int64_t* ckpt_offsets;
    void* scratch;
    int32_t tile_threads;
    int32_t tile_split;
    const int64_t* order;
    int64_t num_increments;
    int32_t tile_expand;  /* register levels of the tile (0: any) */
    int32_t tile_kind;    /* 0: Horner register tile, 2: thread-owned subtree tile */
} rpp_hip_signature_backward_desc;
int64_t rpp_hip_signature_backward_scratch_bytes(const rpp_hip_signature_backward_desc* desc,
                                                 int32_t dtype_bytes);
int rpp_hip_signature_backward_f32(const rpp_hip_signature_backward_desc* desc, void* stream);
int rpp_hip_signature_backward_f64(const rpp_hip_signature_backward_desc* desc, void* stream);

/* The checkpoint spacing the library recommends for a batched signature
 * whose backward pass will follow, for paths of at most max_length
 * increments: 0 (reconstruct every state by exp(-x)) in FP64 and for FP32
 * paths shorter than 32 increments, else ceil(sqrt(max_length)) (sqrt(L)
 * checkpointing).  In FP32 the reconstruction drift roughly doubles the
 * gradient error from 32 increments up; in FP64 it stays below 5e-13.
 * Returns a negative ErrorCode for a dtype other than 4 or 8 bytes or a
 * negative length.  No reference counterpart. */
int32_t rpp_hip_signature_checkpoint_every(int32_t dtype_bytes, int64_t max_length);

/* rpp_hip_signature_backward_* allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid allocation).
 * Same descriptor and results (the same kernels), restricted to what needs no
 * library workspace: the thread-owned subtree tiles (tile_kind 2;
 * tile_threads 0 takes the shape's first such tile), a device queue order
 * (order != NULL, num_increments >= offsets[n]) and scratch == NULL.  The
 * per-row partial sums (num_increments x width^2 values) live in a fixed
 * 64 MiB device array that the library holds on every GPU it runs on until
 * the process ends; larger calls are refused with RPP_HIP_INVALID_ARGUMENT.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams, threads or processes' graphs replayed together) race
 * and must be ordered by the caller. */
int rpp_hip_signature_backward_noalloc_f32(const rpp_hip_signature_backward_desc* desc, void* stream);
int rpp_hip_signature_backward_noalloc_f64(const rpp_hip_signature_backward_desc* desc, void* stream);

/* The compiled tiles of rpp_hip_signature_*: writes up to `capacity`
 * 6-tuples (width, depth, threads, split, kind, expand) for the dtype (4 or 8
 * bytes) and returns how many exist.  kind 0 tiles (level-synchronous Horner,
 * the top `expand` subtree levels in registers and the rest in LDS) serve
 * forward and backward (split <= 1); kind 1 tiles (two barriers per
 * increment, `expand` levels per job) serve the forward pass, and so do kind 2
 * tiles (one thread per (path, prefix of `split` letters), each keeping its
 * subtree and a private copy of its prefix chain in registers: no LDS, no
 * barriers; `threads` is the work-group size; `expand` is 1 when the tile
 * also has a backward).  A descriptor selects one with
 * tile_threads/tile_split/tile_kind/tile_expand; tile_threads = 0 takes the
 * library default for (width, depth), tile_threads < 0 the storage-generic
 * kernel.  No reference counterpart: the reference has no batched signature
 * entry point. */
int rpp_hip_signature_tiles(int32_t dtype_bytes, int32_t* out, int32_t capacity);

/* The tile that serves plain ft_fmexp (degree-one exponent, full output
 * window, unit coefficient stride) for (width, depth) and the dtype: returns
 * 2 for a coalesced subtree tile (k_fmexp_coal), writing its threads and
 * leading letters fixed, 3 for a sibling-group wavefront tile (k_fmexp_wave),
 * writing its threads and 0, or 0 when a generic kernel serves it.  No
 * reference counterpart (introspection of this strategy). */
int rpp_hip_fmexp_tile(int32_t dtype_bytes, int32_t width, int32_t depth, int32_t* threads, int32_t* split);

/* ------------------------------------------------------------------ misc */

/* Message for the last non-zero return on the calling thread. */
const char* rpp_hip_last_error(void);
/* Library version string. */
const char* rpp_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RPP_HIP_H */
