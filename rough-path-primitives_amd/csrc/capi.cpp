// extern "C" entry points of librpp_hip.so (include/rpp_hip.h): descriptor
// validation on the host, then dispatch to the kernels.  Validation follows
// the reference's contract: windows inside [0, depth], out not broadcast,
// fmexp's out distinct from its multiplier (operations/intermediate/
// ft_fmexp.hpp:73-86), Lie windows from degree 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "alias.hpp"
#include "common.hpp"
#include "fmexp_coal.hpp"
#include "ft_ops.hpp"
#include "lie.hpp"
#include "signature.hpp"
#include "sparse.hpp"
#include "signature_reg.hpp"
#include "signature_chain.hpp"
#include "signature_thread.hpp"
#include "shuffle_ops.hpp"
#include "vector_ops.hpp"

namespace rpp_hip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

const char* last_error_cstr() { return g_last_error.c_str(); }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int map_hip_error(hipError_t err) {
    switch (err) {
        case hipSuccess: return RPP_HIP_SUCCESS;
        case hipErrorOutOfMemory: return RPP_HIP_OUT_OF_RESOURCES;
        case hipErrorInvalidValue:
        case hipErrorInvalidDevicePointer:
        case hipErrorInvalidConfiguration:
        case hipErrorInvalidResourceHandle:
        case hipErrorInvalidDevice: return RPP_HIP_INVALID_ARGUMENT;
        case hipErrorNotSupported:
        case hipErrorNoBinaryForGpu:
        case hipErrorInvalidDeviceFunction: return RPP_HIP_NOT_IMPLEMENTED;
        case hipErrorLaunchTimeOut: return RPP_HIP_TIMEOUT;
        case hipErrorLaunchOutOfResources: return RPP_HIP_OUT_OF_RESOURCES;
        default: return RPP_HIP_INTERNAL;
    }
}

int check_hip(hipError_t err, const char* what) {
    if (err == hipSuccess) return 0;
    return fail(map_hip_error(err), std::string(what) + ": " + hipGetErrorString(err));
}

unsigned grid_for(int64_t items, unsigned cap) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(items, cap)));
}

DevBasis make_basis(int width, int depth) {
    DevBasis B{};
    B.width = width;
    B.depth = depth;
    uint32_t L = 0;
    while ((uint64_t{1} << L) < static_cast<uint64_t>(width)) ++L;
    const uint64_t num = uint64_t{1} << (32 + L);
    B.magic = num / static_cast<uint64_t>(width) + ((num % static_cast<uint64_t>(width)) ? 1 : 0);
    B.shift = 32 + L;
    uint64_t p = 1;
    int64_t s = 0;
    for (int k = 0; k <= depth + 1 && k < kMaxDepth + 2; ++k) {
        B.pow[k] = static_cast<uint32_t>(std::min<uint64_t>(p, 0xffffffffu));
        B.start[k] = s;
        s += static_cast<int64_t>(p);
        p *= static_cast<uint64_t>(width);
    }
    return B;
}

namespace {

int check_basis(const rpp_hip_basis& b) {
    if (b.width < 1 || b.width > 65535) return fail(RPP_HIP_INVALID_ARGUMENT, "width must be in [1, 65535]");
    if (b.depth < 0 || b.depth > kMaxDepth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "depth must be in [0, 30]");
    // width^depth, stopped as soon as it reaches 2^31 (so it cannot wrap)
    uint64_t p = 1;
    for (int k = 0; k < b.depth && p < (uint64_t{1} << 31); ++k) p *= static_cast<uint64_t>(b.width);
    if (p >= (uint64_t{1} << 31))
        return fail(RPP_HIP_OUT_OF_BOUNDS, "width^depth must be below 2^31 words per level");
    return 0;
}

// A batch view's largest element offset, (n - 1) batch_stride +
// (end - 1) coeff_stride, must stay below 2^59 elements: then every address
// the kernels form (in bytes of a double) stays below 2^62, and no offset
// arithmetic on the host or the device can wrap.
bool extent_fits(int64_t n, int64_t bs, int64_t cs, int64_t end) {
    if (n <= 0 || end <= 0) return true;
    int64_t a, b, c;
    if (__builtin_mul_overflow(n - 1, bs, &a) || __builtin_mul_overflow(end - 1, cs, &b) ||
        __builtin_add_overflow(a, b, &c))
        return false;
    return c >= 0 && c < (int64_t{1} << 59);
}

int check_view(const rpp_hip_tensor_batch& t, const rpp_hip_basis& b, int64_t n, const char* name,
               bool output) {
    if (t.min_degree < 0 || t.max_degree > b.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": window outside [0, depth]");
    if (t.coeff_stride < 1)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": coeff_stride must be >= 1");
    if (t.batch_stride < 0)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": negative batch_stride");
    if (output && n > 1 && t.batch_stride == 0 && t.min_degree <= t.max_degree)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": an output cannot be broadcast");
    if (n > 0 && t.min_degree <= t.max_degree && t.data == nullptr)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": null data");
    if (t.min_degree <= t.max_degree) {
        // one past the window's last coefficient (width^k < 2^31 per level)
        int64_t end = 0, p = 1;
        for (int k = 0; k <= t.max_degree; ++k, p *= b.width) end += p;
        if (!extent_fits(n, t.batch_stride, t.coeff_stride, end))
            return fail(RPP_HIP_OUT_OF_BOUNDS, std::string(name) + ": the view extends beyond 2^59 elements");
    }
    return 0;
}

int check_count(int64_t n) {
    if (n < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_batches");
    return 0;
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- aliasing

// The elements a tensor batch view's window covers (start: the basis's
// degree offsets).
ViewSet view_set(const rpp_hip_tensor_batch& t, const int64_t* start) {
    if (t.min_degree > t.max_degree) return ViewSet{t.data, t.batch_stride, t.coeff_stride, 0, 0};
    return ViewSet{t.data, t.batch_stride, t.coeff_stride, start[t.min_degree], start[t.max_degree + 1]};
}

bool same_view(const rpp_hip_tensor_batch& a, const rpp_hip_tensor_batch& b) {
    return a.data == b.data && a.batch_stride == b.batch_stride && a.coeff_stride == b.coeff_stride;
}

ViewSet view_set(const rpp_hip_vector_batch& v) {
    return ViewSet{v.data, v.batch_stride, v.coeff_stride, v.begin, v.begin + v.size};
}

template <typename T>
int check_disjoint(const ViewSet& out, const ViewSet& in, int64_t n, const char* what) {
    if (views_overlap(out, in, n, sizeof(T)))
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(what) + " must not share memory");
    return 0;
}

// An output that shares no element with an input.
template <typename T>
int check_disjoint(const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& in, const int64_t* start, int64_t n,
                   const char* what) {
    return check_disjoint<T>(view_set(out, start), view_set(in, start), n, what);
}

// An output that is either the very same batch as an input (data and
// strides: the per-object in-place forms the kernels serve) or shares no
// element with it.
template <typename T>
int check_same_or_disjoint(const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& in, const int64_t* start,
                           int64_t n, const char* what) {
    if (same_view(out, in)) return 0;
    if (views_overlap(view_set(out, start), view_set(in, start), n, sizeof(T)))
        return fail(RPP_HIP_INVALID_ARGUMENT,
                    std::string(what) + " may share memory only as the same batch (same data and strides)");
    return 0;
}

template <typename T>
int ft_mul_impl(const rpp_hip_ft_mul_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->lhs, d->basis, d->num_batches, "lhs", false));
    RPP_TRY(check_view(d->rhs, d->basis, d->num_batches, "rhs", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->lhs, B.start, d->num_batches, "ft_mul: out and lhs"));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->rhs, B.start, d->num_batches, "ft_mul: out and rhs"));
    return launch_product<T>(B, d->num_batches, d->out, d->lhs, d->rhs, nullptr, 0.0, d->beta, as_stream(s));
}

template <typename T>
int ft_fma_impl(const rpp_hip_ft_fma_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->a, d->basis, d->num_batches, "a", false));
    RPP_TRY(check_view(d->b, d->basis, d->num_batches, "b", false));
    RPP_TRY(check_view(d->c, d->basis, d->num_batches, "c", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->a, B.start, d->num_batches, "ft_fma: out and a"));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->b, B.start, d->num_batches, "ft_fma: out and b"));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->c, B.start, d->num_batches, "ft_fma: out and c"));
    return launch_product<T>(B, d->num_batches, d->out, d->b, d->c, &d->a, d->alpha, d->beta, as_stream(s));
}

template <typename T>
int ft_inplace_fma_impl(const rpp_hip_ft_inplace_fma_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->a, d->basis, d->num_batches, "a", true));
    RPP_TRY(check_view(d->b, d->basis, d->num_batches, "b", false));
    RPP_TRY(check_view(d->c, d->basis, d->num_batches, "c", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_same_or_disjoint<T>(d->a, d->b, B.start, d->num_batches, "ft_inplace_fma: a and b"));
    RPP_TRY(check_same_or_disjoint<T>(d->a, d->c, B.start, d->num_batches, "ft_inplace_fma: a and c"));
    switch (d->variant) {
        case 231:  // a <- alpha*a + beta*b(x)c
            return launch_product<T>(B, d->num_batches, d->a, d->b, d->c, &d->a, d->alpha, d->beta,
                                     as_stream(s));
        case 123:  // a <- alpha*c + beta*a(x)b
            return launch_product<T>(B, d->num_batches, d->a, d->a, d->b, &d->c, d->alpha, d->beta,
                                     as_stream(s));
        case 213:  // a <- alpha*c + beta*b(x)a
            return launch_product<T>(B, d->num_batches, d->a, d->b, d->a, &d->c, d->alpha, d->beta,
                                     as_stream(s));
        default: return fail(RPP_HIP_INVALID_ARGUMENT, "variant must be 231, 123 or 213");
    }
}

template <typename T>
int ft_inplace_mul_impl(const rpp_hip_ft_inplace_mul_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->lhs, d->basis, d->num_batches, "lhs", true));
    RPP_TRY(check_view(d->rhs, d->basis, d->num_batches, "rhs", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_same_or_disjoint<T>(d->lhs, d->rhs, B.start, d->num_batches, "ft_inplace_mul: lhs and rhs"));
    return launch_product<T>(B, d->num_batches, d->lhs, d->lhs, d->rhs, nullptr, 0.0, d->beta, as_stream(s));
}

template <typename T>
int unary_impl(int op, const rpp_hip_ft_unary_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->arg, d->basis, d->num_batches, "arg", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_disjoint<T>(d->out, d->arg, B.start, d->num_batches, "out and arg"));
    return launch_horner<T>(op, B, d->num_batches, d->out, d->arg, nullptr, as_stream(s));
}

template <typename T>
int antipode_impl(bool sign_by_degree, const rpp_hip_ft_unary_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->arg, d->basis, d->num_batches, "arg", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_disjoint<T>(d->out, d->arg, B.start, d->num_batches, "out and arg"));
    return launch_antipode<T>(sign_by_degree, B, d->num_batches, d->out, d->arg, as_stream(s));
}

template <typename T>
int fmexp_impl(const rpp_hip_ft_fmexp_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->multiplier, d->basis, d->num_batches, "multiplier", false));
    RPP_TRY(check_view(d->exponent, d->basis, d->num_batches, "exponent", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_disjoint<T>(d->out, d->multiplier, B.start, d->num_batches, "ft_fmexp: out and the multiplier"));
    RPP_TRY(check_disjoint<T>(d->out, d->exponent, B.start, d->num_batches, "ft_fmexp: out and the exponent"));
    return launch_horner<T>(kOpFmexp, B, d->num_batches, d->out, d->exponent, &d->multiplier, as_stream(s));
}

template <typename T>
int log_adjoint_impl(const rpp_hip_ft_log_adjoint_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    const int64_t n = d->num_batches;
    if (n < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_batches");
    if (n == 0) return 0;
    RPP_TRY(check_view(d->arg, d->basis, n, "arg", false));
    RPP_TRY(check_view(d->cotangent, d->basis, n, "cotangent", false));
    RPP_TRY(check_view(d->grad_arg, d->basis, n, "grad_arg", true));
    if (d->grad_arg.min_degree != d->arg.min_degree || d->grad_arg.max_degree != d->arg.max_degree)
        return fail(RPP_HIP_INVALID_ARGUMENT, "ft_log adjoint: grad_arg must have arg's window");
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_disjoint<T>(d->grad_arg, d->arg, B.start, n, "ft_log adjoint: grad_arg and arg"));
    RPP_TRY(check_disjoint<T>(d->grad_arg, d->cotangent, B.start, n, "ft_log adjoint: grad_arg and the cotangent"));
    return launch_log_adjoint<T>(B, n, d->grad_arg, d->arg, d->cotangent, as_stream(s));
}

template <typename T>
int pairing_impl(const rpp_hip_tensor_pairing_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_count(d->num_batches));
    RPP_TRY(check_view(d->functional, d->basis, d->num_batches, "functional", false));
    RPP_TRY(check_view(d->arg, d->basis, d->num_batches, "arg", false));
    if (d->num_batches > 0 && d->out.data == nullptr) return fail(RPP_HIP_INVALID_ARGUMENT, "out: null data");
    if (d->out.stride < 0 || (d->num_batches > 1 && d->out.stride == 0))
        return fail(RPP_HIP_INVALID_ARGUMENT, "out: the stride must be >= 1 (an output cannot be broadcast)");
    if (!extent_fits(d->num_batches, d->out.stride, 1, 1))
        return fail(RPP_HIP_OUT_OF_BOUNDS, "out: the batch extends beyond 2^59 elements");
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    const ViewSet out{d->out.data, d->out.stride, 1, 0, 1};
    RPP_TRY(check_disjoint<T>(out, view_set(d->functional, B.start), d->num_batches, "tensor_pairing: out and the functional"));
    RPP_TRY(check_disjoint<T>(out, view_set(d->arg, B.start), d->num_batches, "tensor_pairing: out and arg"));
    return launch_pairing<T>(B, d->num_batches, d->out, d->functional, d->arg, as_stream(s));
}

template <typename T>
int adjoint_impl(bool left, const rpp_hip_ft_adjoint_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    RPP_TRY(check_view(d->out, d->basis, d->num_batches, "out", true));
    RPP_TRY(check_view(d->op, d->basis, d->num_batches, "op", false));
    RPP_TRY(check_view(d->arg, d->basis, d->num_batches, "arg", false));
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    RPP_TRY(check_disjoint<T>(d->out, d->op, B.start, d->num_batches, "adjoint multiplication: out and op"));
    RPP_TRY(check_disjoint<T>(d->out, d->arg, B.start, d->num_batches, "adjoint multiplication: out and arg"));
    return launch_adjoint<T>(left, B, d->num_batches, d->out, d->op, d->arg, as_stream(s));
}

int check_vec(const rpp_hip_vector_batch& v, int64_t n, const char* name) {
    if (v.begin < 0 || v.size < 0 || v.coeff_stride < 1 || v.batch_stride < 0)
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": invalid vector batch");
    int64_t end;
    if (__builtin_add_overflow(v.begin, v.size, &end) || !extent_fits(n, v.batch_stride, v.coeff_stride, end))
        return fail(RPP_HIP_OUT_OF_BOUNDS, std::string(name) + ": the view extends beyond 2^59 elements");
    return 0;
}

template <typename T>
int spmv_impl(const rpp_hip_spmv_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_count(d->num_batches));
    RPP_TRY(check_vec(d->out, d->num_batches, "out"));
    RPP_TRY(check_vec(d->arg, d->num_batches, "arg"));
    const auto& M = d->matrix;
    if (M.format != RPP_HIP_CSR && M.format != RPP_HIP_CSC)
        return fail(RPP_HIP_INVALID_ARGUMENT, "matrix format must be CSR or CSC");
    if (M.rows < 0 || M.cols < 0 || M.nnz < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative matrix shape");
    if (d->num_batches > 0 && (M.offsets == nullptr || (M.nnz > 0 && (!M.values || !M.indices))))
        return fail(RPP_HIP_INVALID_ARGUMENT, "null matrix arrays");
    if (d->num_batches == 0) return 0;
    RPP_TRY(check_disjoint<T>(view_set(d->out), view_set(d->arg), d->num_batches,
                              "sparse_matrix_vector_product: out and arg"));
    PreparedMatrix P;
    RPP_TRY(prepare_matrix<T>(M, &P, as_stream(s)));
    return launch_spmv<T>(d->num_batches, d->out, d->arg, P.csr, d->alpha, as_stream(s), 0, P.csc);
}

// tensor_to_lie / lie_to_tensor with the library's matrices.  Both are block
// diagonal by degree (a Lie element of degree k expands into, and is
// recovered from, tensor words of degree k only), so a window pair converts
// the degrees [dlo, dhi] that both windows hold, reading the argument only
// there, and the rest of the output window is 0 (the argument is zero outside
// its window).  With the reference's windows (tensor from degree 0, Lie from
// degree 1) this is exactly its view-relative product
// (cpu/.../linalg/sparse_matrix_vector.hpp:58-59, 148-149).
template <typename T>
int lie_convert_impl(bool to_lie, const rpp_hip_lie_convert_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    if (d->basis.width < 2) return fail(RPP_HIP_NOT_IMPLEMENTED, "Lie conversions need width >= 2");
    RPP_TRY(check_count(d->num_batches));
    RPP_TRY(check_view(d->tensor, d->basis, d->num_batches, "tensor", !to_lie));
    if (d->lie.min_degree < 1 || d->lie.max_degree > d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "lie: window must lie in [1, depth]");
    RPP_TRY(check_view(d->lie, d->basis, d->num_batches, "lie", to_lie));
    const int64_t n = d->num_batches;
    auto md = lie_metadata(d->basis.width, d->basis.depth);
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    const std::vector<int64_t>& lb = md->degree_begin;  // Lie storage index of degree k's first key
    const rpp_hip_tensor_batch& out = to_lie ? d->lie : d->tensor;
    const rpp_hip_tensor_batch& arg = to_lie ? d->tensor : d->lie;
    auto ostart = [&](int k) { return to_lie ? lb[static_cast<size_t>(k)] : B.start[k]; };
    auto astart = [&](int k) { return to_lie ? B.start[k] : lb[static_cast<size_t>(k)]; };
    {
        auto set = [&](const rpp_hip_tensor_batch& t, bool lie) {
            if (t.min_degree > t.max_degree) return ViewSet{t.data, t.batch_stride, t.coeff_stride, 0, 0};
            const int64_t b0 = lie ? lb[static_cast<size_t>(t.min_degree)] : B.start[t.min_degree];
            const int64_t b1 = lie ? lb[static_cast<size_t>(t.max_degree) + 1] : B.start[t.max_degree + 1];
            return ViewSet{t.data, t.batch_stride, t.coeff_stride, b0, b1};
        };
        RPP_TRY(check_disjoint<T>(set(d->tensor, false), set(d->lie, true), n, "Lie conversion: tensor and lie"));
    }
    const DeviceCsr* l2t = nullptr;
    const DeviceCsr* t2l = nullptr;
    RPP_TRY(device_lie_matrices<T>(d->basis.width, d->basis.depth, &l2t, &t2l));
    const int dlo = std::max({d->tensor.min_degree, d->lie.min_degree, 1});
    const int dhi = std::min(d->tensor.max_degree, d->lie.max_degree);
    const hipStream_t stream = as_stream(s);
    if (out.min_degree <= out.max_degree && n > 0) {
        const int64_t w0 = ostart(out.min_degree), w1 = ostart(out.max_degree + 1);
        const int64_t c0 = dlo <= dhi ? ostart(dlo) : w1, c1 = dlo <= dhi ? ostart(dhi + 1) : w1;
        if (c0 > w0)
            RPP_TRY(launch_vector_op<T>(kVecSetConstant, n, out, IndexRange{w0, c0}, nullptr, IndexRange{0, 0},
                                        nullptr, IndexRange{0, 0}, 0.0, 0.0, stream));
        if (w1 > c1)
            RPP_TRY(launch_vector_op<T>(kVecSetConstant, n, out, IndexRange{c1, w1}, nullptr, IndexRange{0, 0},
                                        nullptr, IndexRange{0, 0}, 0.0, 0.0, stream));
    }
    if (dlo > dhi || n == 0) return 0;
    const DeviceCsr* m = to_lie ? t2l : l2t;
    // the degrees [lo, hi] in one launch: matrix rows are Lie keys - 1 (T2L)
    // or tensor words (L2T), columns the other way round
    auto convert = [&](int lo, int hi) {
        const int64_t obeg = ostart(lo), rows = ostart(hi + 1) - obeg;
        const int64_t abeg = astart(lo), ncols = astart(hi + 1) - abeg;
        const int64_t row0 = to_lie ? obeg - 1 : obeg;
        const int64_t col0 = to_lie ? abeg : abeg - 1;
        const rpp_hip_vector_batch ov{out.data, out.batch_stride, out.coeff_stride, obeg, rows};
        const rpp_hip_vector_batch av{arg.data, arg.batch_stride, arg.coeff_stride, abeg, ncols};
        if (lie_spmv_fits<T>(*m, av)) return launch_lie_spmv<T>(n, ov, av, *m, row0, rows, col0, stream);
        const rpp_hip_sparse_matrix M{RPP_HIP_CSR, m->values, m->indices, m->offsets + row0, m->nnz, rows, ncols};
        return launch_spmv<T>(n, ov, av, M, 1.0, stream, col0);
    };
    return convert(dlo, dhi);
}

// tensor_to_lie / lie_to_tensor with the caller's matrix, as the reference's
// operations take it: the view-relative product out = M arg
// (operations/basic/tensor_to_lie.hpp:53-60, lie_to_tensor.hpp:57-65).  With
// the reference's windows (tensor [0, depth], Lie [1, depth]) and a matrix
// equal to the library's Hall T2L (L2T) of the basis, row for row in column
// order, the product is the library's Lie conversion, which sums each row in
// that order: the same bits.  Any other matrix or window pair takes the
// deterministic sparse product.
template <typename T>
int lie_matrix_convert_impl(bool to_lie, const rpp_hip_lie_matrix_convert_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    if (d->basis.width < 2) return fail(RPP_HIP_NOT_IMPLEMENTED, "Lie conversions need width >= 2");
    RPP_TRY(check_count(d->num_batches));
    RPP_TRY(check_view(d->tensor, d->basis, d->num_batches, "tensor", !to_lie));
    if (d->lie.min_degree < 1 || d->lie.max_degree > d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "lie: window must lie in [1, depth]");
    RPP_TRY(check_view(d->lie, d->basis, d->num_batches, "lie", to_lie));
    const auto& M = d->matrix;
    if (M.format != RPP_HIP_CSR && M.format != RPP_HIP_CSC)
        return fail(RPP_HIP_INVALID_ARGUMENT, "matrix format must be CSR or CSC");
    if (M.rows < 0 || M.cols < 0 || M.nnz < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative matrix shape");
    if (d->num_batches > 0 && (M.offsets == nullptr || (M.nnz > 0 && (!M.values || !M.indices))))
        return fail(RPP_HIP_INVALID_ARGUMENT, "null matrix arrays");
    if (d->num_batches == 0) return 0;
    const hipStream_t stream = as_stream(s);
    const int depth = d->basis.depth;
    if (d->tensor.min_degree == 0 && d->tensor.max_degree == depth && d->lie.min_degree == 1 &&
        d->lie.max_degree == depth) {
        bool hall = false;
        RPP_TRY(matrix_is_hall<T>(M, d->basis.width, depth, to_lie, &hall, stream));
        if (hall) {
            const rpp_hip_lie_convert_desc c{d->basis, d->num_batches, d->tensor, d->lie};
            return lie_convert_impl<T>(to_lie, &c, s);
        }
    }
    // view-relative vectors: the tensor and Lie windows as index ranges
    auto md = lie_metadata(d->basis.width, depth);
    const DevBasis B = make_basis(d->basis.width, depth);
    auto tensor_vec = [&](const rpp_hip_tensor_batch& t) {
        return rpp_hip_vector_batch{t.data, t.batch_stride, t.coeff_stride, B.start[t.min_degree],
                                    B.start[t.max_degree + 1] - B.start[t.min_degree]};
    };
    auto lie_vec = [&](const rpp_hip_tensor_batch& l) {
        const auto& lb = md->degree_begin;
        return rpp_hip_vector_batch{l.data, l.batch_stride, l.coeff_stride, lb[static_cast<size_t>(l.min_degree)],
                                    lb[static_cast<size_t>(l.max_degree) + 1] - lb[static_cast<size_t>(l.min_degree)]};
    };
    const rpp_hip_vector_batch out = to_lie ? lie_vec(d->lie) : tensor_vec(d->tensor);
    const rpp_hip_vector_batch arg = to_lie ? tensor_vec(d->tensor) : lie_vec(d->lie);
    if (M.rows > out.size || M.cols > arg.size)
        return fail(RPP_HIP_INVALID_ARGUMENT, "matrix larger than the views it maps between");
    RPP_TRY(check_disjoint<T>(view_set(out), view_set(arg), d->num_batches, "Lie conversion: tensor and lie"));
    PreparedMatrix P;
    RPP_TRY(prepare_matrix<T>(M, &P, stream));
    return launch_spmv<T>(d->num_batches, out, arg, P.csr, 1.0, stream, 0, P.csc);
}

// ------------------------------------------------------- the shuffle family

int check_shuffle_basis(const rpp_hip_basis& b) {
    RPP_TRY(check_basis(b));
    if (b.depth > kShuffleMaxDepth)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "shuffle products are limited to depth 16");
    return 0;
}

template <typename T>
int shuffle_impl(int op, const rpp_hip_basis& basis, int64_t n, const rpp_hip_tensor_batch& out,
                 const rpp_hip_tensor_batch& p, const rpp_hip_tensor_batch& q, const rpp_hip_tensor_batch* r,
                 double alpha, double beta, void* s) {
    RPP_TRY(check_shuffle_basis(basis));
    if (n < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_batches");
    RPP_TRY(check_view(out, basis, n, "out", true));
    RPP_TRY(check_view(p, basis, n, "operand", false));
    RPP_TRY(check_view(q, basis, n, "operand", false));
    if (r) RPP_TRY(check_view(*r, basis, n, "addend", false));
    const DevBasis B = make_basis(basis.width, basis.depth);
    RPP_TRY(check_disjoint<T>(out, p, B.start, n, "shuffle products: the output and a factor"));
    RPP_TRY(check_disjoint<T>(out, q, B.start, n, "shuffle products: the output and a factor"));
    if (r) RPP_TRY(check_same_or_disjoint<T>(out, *r, B.start, n, "st_fma: out and the addend"));
    return launch_shuffle<T>(op, B, n, out, p, q, r, alpha, beta, as_stream(s));
}

template <typename T>
int st_mul_impl(const rpp_hip_ft_mul_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    return shuffle_impl<T>(kStMul, d->basis, d->num_batches, d->out, d->lhs, d->rhs, nullptr, 0.0, d->beta, s);
}

template <typename T>
int st_fma_impl(const rpp_hip_ft_fma_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    return shuffle_impl<T>(kStFma, d->basis, d->num_batches, d->out, d->b, d->c, &d->a, d->alpha, d->beta, s);
}

template <typename T>
int st_inplace_fma_impl(const rpp_hip_st_inplace_fma_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    return shuffle_impl<T>(kStInplace, d->basis, d->num_batches, d->a, d->b, d->c, nullptr, d->alpha, d->beta, s);
}

template <typename T>
int st_adj_mul_impl(const rpp_hip_ft_adjoint_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    return shuffle_impl<T>(kStAdjMul, d->basis, d->num_batches, d->out, d->op, d->arg, nullptr, 0.0, 1.0, s);
}

template <typename T, int OP>
int half_shuffle_impl(const rpp_hip_half_shuffle_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    return shuffle_impl<T>(OP, d->basis, d->num_batches, d->out, d->integrator, d->integrand, nullptr, 0.0, d->beta,
                           s);
}

template <typename T>
int left_hs_mul_impl(const rpp_hip_half_shuffle_desc* d, void* s) { return half_shuffle_impl<T, kHsMul>(d, s); }
template <typename T>
int left_hs_adj_lmul_impl(const rpp_hip_half_shuffle_desc* d, void* s) { return half_shuffle_impl<T, kHsAdjL>(d, s); }
template <typename T>
int left_hs_adj_rmul_impl(const rpp_hip_half_shuffle_desc* d, void* s) { return half_shuffle_impl<T, kHsAdjR>(d, s); }

// ---------------------------------------------------------- linear operations

// Number of Lie elements of degree k over `width` letters (Witt's formula:
// (1/k) sum_{d | k} mu(d) width^(k/d)), the size of degree k of the Hall-set
// LieBasis (basis/hall_basis.hpp:64-88).
int64_t witt_dimension(int width, int k) {
    int64_t sum = 0;
    for (int d = 1; d <= k; ++d) {
        if (k % d) continue;
        int mu = 1, m = d;
        for (int p = 2; p * p <= m; ++p) {
            if (m % p) continue;
            m /= p;
            if (m % p == 0) {
                mu = 0;
                break;
            }
            mu = -mu;
        }
        if (mu != 0 && m > 1) mu = -mu;
        if (mu == 0) continue;
        int64_t pw = 1;
        for (int e = 0; e < k / d; ++e) pw *= width;
        sum += mu * pw;
    }
    return sum / k;
}

// Degree offsets start_of_degree(k), k = 0..depth+1, of a graded basis.
int graded_offsets(const rpp_hip_graded_basis& g, int64_t* start) {
    const rpp_hip_basis b{g.width, g.depth};
    RPP_TRY(check_basis(b));
    if (g.kind == RPP_HIP_TENSOR_BASIS) {
        const DevBasis B = make_basis(g.width, g.depth);
        for (int k = 0; k <= g.depth + 1; ++k) start[k] = B.start[k];
        return 0;
    }
    if (g.kind != RPP_HIP_LIE_BASIS) return fail(RPP_HIP_INVALID_ARGUMENT, "basis kind must be tensor or Lie");
    start[0] = 0;
    start[1] = 1;
    for (int k = 1; k <= g.depth; ++k) start[k + 1] = start[k] + witt_dimension(g.width, k);
    return 0;
}

IndexRange range_of(const int64_t* start, const rpp_hip_tensor_batch& t) {
    if (t.min_degree > t.max_degree) return IndexRange{0, 0};
    return IndexRange{start[t.min_degree], start[t.max_degree + 1]};
}

IndexRange overlap(IndexRange a, IndexRange b) {
    return IndexRange{std::max(a.begin, b.begin), std::min(a.end, b.end)};
}

bool nonempty(const rpp_hip_tensor_batch& t) { return t.min_degree <= t.max_degree; }

template <typename T>
int vector_assign_impl(const rpp_hip_vector_assign_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    int64_t start[kMaxDepth + 2];
    RPP_TRY(graded_offsets(d->basis, start));
    RPP_TRY(check_count(d->num_batches));
    const rpp_hip_basis b{d->basis.width, d->basis.depth};
    RPP_TRY(check_view(d->out, b, d->num_batches, "out", true));
    RPP_TRY(check_view(d->arg, b, d->num_batches, "arg", false));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->arg, start, d->num_batches, "vector_assign: out and arg"));
    const IndexRange o = range_of(start, d->out);
    return launch_vector_op<T>(kVecAssign, d->num_batches, d->out, o, &d->arg, overlap(o, range_of(start, d->arg)),
                               nullptr, IndexRange{0, 0}, 0.0, 0.0, as_stream(s));
}

template <typename T>
int vector_add_impl(const rpp_hip_vector_add_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    int64_t start[kMaxDepth + 2];
    RPP_TRY(graded_offsets(d->basis, start));
    RPP_TRY(check_count(d->num_batches));
    const rpp_hip_basis b{d->basis.width, d->basis.depth};
    RPP_TRY(check_view(d->out, b, d->num_batches, "out", true));
    RPP_TRY(check_view(d->lhs, b, d->num_batches, "lhs", false));
    RPP_TRY(check_view(d->rhs, b, d->num_batches, "rhs", false));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->lhs, start, d->num_batches, "vector_add: out and lhs"));
    RPP_TRY(check_same_or_disjoint<T>(d->out, d->rhs, start, d->num_batches, "vector_add: out and rhs"));
    return launch_vector_op<T>(kVecAdd, d->num_batches, d->out, range_of(start, d->out), &d->lhs,
                               range_of(start, d->lhs), &d->rhs, range_of(start, d->rhs), d->alpha, d->beta,
                               as_stream(s));
}

template <typename T>
int vector_inplace_add_impl(const rpp_hip_vector_inplace_add_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    int64_t start[kMaxDepth + 2];
    RPP_TRY(graded_offsets(d->basis, start));
    RPP_TRY(check_count(d->num_batches));
    const rpp_hip_basis b{d->basis.width, d->basis.depth};
    RPP_TRY(check_view(d->lhs, b, d->num_batches, "lhs", true));
    RPP_TRY(check_view(d->rhs, b, d->num_batches, "rhs", false));
    RPP_TRY(check_same_or_disjoint<T>(d->lhs, d->rhs, start, d->num_batches, "vector_inplace_add: lhs and rhs"));
    if (!nonempty(d->lhs) || !nonempty(d->rhs)) return 0;
    const IndexRange o = overlap(range_of(start, d->lhs), range_of(start, d->rhs));
    return launch_vector_op<T>(kVecInplaceAdd, d->num_batches, d->lhs, o, &d->rhs, o, nullptr, IndexRange{0, 0},
                               d->alpha, 0.0, as_stream(s));
}

template <typename T>
int vector_scalar_impl(int op, const rpp_hip_vector_scalar_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    int64_t start[kMaxDepth + 2];
    RPP_TRY(graded_offsets(d->basis, start));
    RPP_TRY(check_count(d->num_batches));
    const rpp_hip_basis b{d->basis.width, d->basis.depth};
    RPP_TRY(check_view(d->vec, b, d->num_batches, "vec", true));
    return launch_vector_op<T>(op, d->num_batches, d->vec, range_of(start, d->vec), nullptr, IndexRange{0, 0},
                               nullptr, IndexRange{0, 0}, d->scalar, 0.0, as_stream(s));
}

}  // namespace
}  // namespace rpp_hip

using namespace rpp_hip;

extern "C" {

#define RPP_C_ENTRY(NAME, DESC, CALL)                                             \
    int NAME##_f32(const DESC* d, void* s) { return CALL<float>(d, s); }          \
    int NAME##_f64(const DESC* d, void* s) { return CALL<double>(d, s); }

RPP_C_ENTRY(rpp_hip_ft_mul, rpp_hip_ft_mul_desc, ft_mul_impl)
RPP_C_ENTRY(rpp_hip_ft_fma, rpp_hip_ft_fma_desc, ft_fma_impl)
RPP_C_ENTRY(rpp_hip_ft_inplace_fma, rpp_hip_ft_inplace_fma_desc, ft_inplace_fma_impl)
RPP_C_ENTRY(rpp_hip_ft_inplace_mul, rpp_hip_ft_inplace_mul_desc, ft_inplace_mul_impl)
RPP_C_ENTRY(rpp_hip_ft_fmexp, rpp_hip_ft_fmexp_desc, fmexp_impl)
RPP_C_ENTRY(rpp_hip_tensor_pairing, rpp_hip_tensor_pairing_desc, pairing_impl)
RPP_C_ENTRY(rpp_hip_sparse_matrix_vector_product, rpp_hip_spmv_desc, spmv_impl)
RPP_C_ENTRY(rpp_hip_st_mul, rpp_hip_ft_mul_desc, st_mul_impl)
RPP_C_ENTRY(rpp_hip_st_fma, rpp_hip_ft_fma_desc, st_fma_impl)
RPP_C_ENTRY(rpp_hip_st_inplace_fma, rpp_hip_st_inplace_fma_desc, st_inplace_fma_impl)
RPP_C_ENTRY(rpp_hip_st_adj_mul, rpp_hip_ft_adjoint_desc, st_adj_mul_impl)
RPP_C_ENTRY(rpp_hip_left_hs_mul, rpp_hip_half_shuffle_desc, left_hs_mul_impl)
RPP_C_ENTRY(rpp_hip_left_hs_adj_lmul, rpp_hip_half_shuffle_desc, left_hs_adj_lmul_impl)
RPP_C_ENTRY(rpp_hip_left_hs_adj_rmul, rpp_hip_half_shuffle_desc, left_hs_adj_rmul_impl)
RPP_C_ENTRY(rpp_hip_vector_assign, rpp_hip_vector_assign_desc, vector_assign_impl)
RPP_C_ENTRY(rpp_hip_vector_add, rpp_hip_vector_add_desc, vector_add_impl)
RPP_C_ENTRY(rpp_hip_vector_inplace_add, rpp_hip_vector_inplace_add_desc, vector_inplace_add_impl)

#define RPP_C_SCALAR_ENTRY(NAME, OP)                                                                     \
    int NAME##_f32(const rpp_hip_vector_scalar_desc* d, void* s) { return vector_scalar_impl<float>(OP, d, s); } \
    int NAME##_f64(const rpp_hip_vector_scalar_desc* d, void* s) { return vector_scalar_impl<double>(OP, d, s); }

RPP_C_SCALAR_ENTRY(rpp_hip_vector_scalar_multiply, kVecScale)
RPP_C_SCALAR_ENTRY(rpp_hip_vector_set_constant, kVecSetConstant)
RPP_C_SCALAR_ENTRY(rpp_hip_tensor_set_identity, kVecSetIdentity)
RPP_C_SCALAR_ENTRY(rpp_hip_tensor_add_identity, kVecAddIdentity)
#undef RPP_C_SCALAR_ENTRY

int rpp_hip_tensor_antipode_f32(const rpp_hip_ft_unary_desc* d, void* s) { return antipode_impl<float>(true, d, s); }
int rpp_hip_tensor_antipode_f64(const rpp_hip_ft_unary_desc* d, void* s) {
    return antipode_impl<double>(true, d, s);
}
int rpp_hip_tensor_reflect_f32(const rpp_hip_ft_unary_desc* d, void* s) { return antipode_impl<float>(false, d, s); }
int rpp_hip_tensor_reflect_f64(const rpp_hip_ft_unary_desc* d, void* s) {
    return antipode_impl<double>(false, d, s);
}
int rpp_hip_ft_exp_f32(const rpp_hip_ft_unary_desc* d, void* s) { return unary_impl<float>(kOpExp, d, s); }
int rpp_hip_ft_exp_f64(const rpp_hip_ft_unary_desc* d, void* s) { return unary_impl<double>(kOpExp, d, s); }
int rpp_hip_ft_log_f32(const rpp_hip_ft_unary_desc* d, void* s) { return unary_impl<float>(kOpLog, d, s); }
int rpp_hip_ft_log_f64(const rpp_hip_ft_unary_desc* d, void* s) { return unary_impl<double>(kOpLog, d, s); }

int rpp_hip_ft_log_adjoint_f32(const rpp_hip_ft_log_adjoint_desc* d, void* s) { return log_adjoint_impl<float>(d, s); }
int rpp_hip_ft_log_adjoint_f64(const rpp_hip_ft_log_adjoint_desc* d, void* s) {
    return log_adjoint_impl<double>(d, s);
}
int rpp_hip_ft_adj_lmul_f32(const rpp_hip_ft_adjoint_desc* d, void* s) { return adjoint_impl<float>(true, d, s); }
int rpp_hip_ft_adj_lmul_f64(const rpp_hip_ft_adjoint_desc* d, void* s) { return adjoint_impl<double>(true, d, s); }
int rpp_hip_ft_adj_rmul_f32(const rpp_hip_ft_adjoint_desc* d, void* s) { return adjoint_impl<float>(false, d, s); }
int rpp_hip_ft_adj_rmul_f64(const rpp_hip_ft_adjoint_desc* d, void* s) { return adjoint_impl<double>(false, d, s); }

int rpp_hip_tensor_to_lie_f32(const rpp_hip_lie_convert_desc* d, void* s) {
    return lie_convert_impl<float>(true, d, s);
}
int rpp_hip_tensor_to_lie_f64(const rpp_hip_lie_convert_desc* d, void* s) {
    return lie_convert_impl<double>(true, d, s);
}
int rpp_hip_lie_to_tensor_f32(const rpp_hip_lie_convert_desc* d, void* s) {
    return lie_convert_impl<float>(false, d, s);
}
int rpp_hip_lie_to_tensor_f64(const rpp_hip_lie_convert_desc* d, void* s) {
    return lie_convert_impl<double>(false, d, s);
}
int rpp_hip_tensor_to_lie_matrix_f32(const rpp_hip_lie_matrix_convert_desc* d, void* s) {
    return lie_matrix_convert_impl<float>(true, d, s);
}
int rpp_hip_tensor_to_lie_matrix_f64(const rpp_hip_lie_matrix_convert_desc* d, void* s) {
    return lie_matrix_convert_impl<double>(true, d, s);
}
int rpp_hip_lie_to_tensor_matrix_f32(const rpp_hip_lie_matrix_convert_desc* d, void* s) {
    return lie_matrix_convert_impl<float>(false, d, s);
}
int rpp_hip_lie_to_tensor_matrix_f64(const rpp_hip_lie_matrix_convert_desc* d, void* s) {
    return lie_matrix_convert_impl<double>(false, d, s);
}
int rpp_hip_sparse_matrix_forget(const void* values) {
    if (!values) return fail(RPP_HIP_INVALID_ARGUMENT, "null values pointer");
    return forget_matrix(values);
}

int rpp_hip_lie_basis_sizes(int32_t width, int32_t depth, int64_t* sizes) {
    const rpp_hip_basis b{width, depth};
    if (int rc = check_basis(b)) return rc;
    if (!sizes) return fail(RPP_HIP_INVALID_ARGUMENT, "null sizes");
    auto md = lie_metadata(width, depth);
    sizes[0] = md->true_size;
    sizes[1] = static_cast<int64_t>(md->l2t_csc_values.size());
    sizes[2] = static_cast<int64_t>(md->t2l_csc_values.size());
    sizes[3] = md->tensor_dim;
    return 0;
}

int rpp_hip_lie_basis_copy(int32_t width, int32_t depth, int64_t* pairs, int64_t* degree_begin,
                           double* l2t_values, int64_t* l2t_indices, int64_t* l2t_offsets,
                           double* t2l_values, int64_t* t2l_indices, int64_t* t2l_offsets) {
    const rpp_hip_basis b{width, depth};
    if (int rc = check_basis(b)) return rc;
    auto md = lie_metadata(width, depth);
    auto cp = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(pairs, md->pairs);
    cp(degree_begin, md->degree_begin);
    cp(l2t_values, md->l2t_csc_values);
    cp(l2t_indices, md->l2t_csc_indices);
    cp(l2t_offsets, md->l2t_csc_offsets);
    cp(t2l_values, md->t2l_csc_values);
    cp(t2l_indices, md->t2l_csc_indices);
    cp(t2l_offsets, md->t2l_csc_offsets);
    return 0;
}

int rpp_hip_lie_basis_save(int32_t width, int32_t depth, const char* path) {
    const rpp_hip_basis b{width, depth};
    if (int rc = check_basis(b)) return rc;
    if (!path || !*path) return fail(RPP_HIP_INVALID_ARGUMENT, "null or empty path");
    auto md = lie_metadata(width, depth);
    std::string err;
    if (!save_lie_metadata(*md, path, &err)) return fail(RPP_HIP_INVALID_ARGUMENT, err);
    return 0;
}

int rpp_hip_lie_basis_load(const char* path, int32_t* width, int32_t* depth) {
    if (!path || !*path) return fail(RPP_HIP_INVALID_ARGUMENT, "null or empty path");
    if (std::FILE* f = std::fopen(path, "rb")) {
        std::fclose(f);
    } else {
        return fail(RPP_HIP_INVALID_ARGUMENT, std::string("cannot open ") + path);
    }
    std::string err;
    auto md = load_lie_metadata(path, &err);
    if (!md) return fail(RPP_HIP_CONTRACT_VIOLATION, err);
    const rpp_hip_basis b{md->width, md->depth};
    if (int rc = check_basis(b)) return rc;
    if (int rc = install_lie_metadata(md, &err)) return fail(rc, err);
    if (width) *width = md->width;
    if (depth) *depth = md->depth;
    return 0;
}

const char* rpp_hip_last_error(void) { return last_error_cstr(); }

const char* rpp_hip_version(void) { return "rpp_hip 0.1.0 (gfx950)"; }

}  // extern "C"

// ------------------------------------------------------- batched signatures

namespace rpp_hip {
namespace {

// The compiled register tile for (width, depth): the requested one
// (threads > 0), else the first listed; backward tiles fix at most one letter.
template <typename T>
const RegTile<T>* pick_tile(const rpp_hip_basis& b, int threads, int split, int reg_levels, bool backward) {
    for (const auto& t : reg_tiles<T>()) {
        if (t.width != b.width || t.depth != b.depth) continue;
        if (backward && (t.backward == nullptr || t.split > 1)) continue;
        if (threads > 0 && (t.threads != threads || t.split != split)) continue;
        if (threads > 0 && reg_levels > 0 && t.reg_levels != reg_levels) continue;
        return &t;
    }
    return nullptr;
}

template <typename T>
int signature_impl(const rpp_hip_signature_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    if (d->num_paths < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_paths");
    if (d->num_paths == 0) return 0;
    RPP_TRY(check_view(d->out, d->basis, d->num_paths, "out", true));
    if (d->out.min_degree != 0 || d->out.max_degree != d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "signature: out must hold the full window [0, depth]");
    if (!d->increments || !d->offsets) return fail(RPP_HIP_INVALID_ARGUMENT, "null increments or offsets");
    if (d->checkpoint_every < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative checkpoint_every");
    if (d->checkpoint_every > 0 && (!d->checkpoints || !d->ckpt_offsets))
        return fail(RPP_HIP_INVALID_ARGUMENT, "checkpoints requested without buffers");
    if (d->basis.width > 256) return fail(RPP_HIP_NOT_IMPLEMENTED, "signature supports width <= 256");
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    hipStream_t stream = as_stream(s);
    const int64_t* order = d->order;
    if (order == nullptr) {
        std::vector<int64_t> lengths;
        int64_t* host_order = nullptr;
        RPP_TRY(path_order(d->offsets, d->num_paths, lengths, &host_order, stream));
        order = host_order;
    }
    if (d->tile_threads < 0) return signature_forward_generic<T>(B, *d, order, stream);
    if (d->tile_kind == 1) {
        for (const auto& c : chain_tiles<T>()) {
            if (c.width != d->basis.width || c.depth != d->basis.depth) continue;
            if (d->tile_threads > 0 && (c.threads != d->tile_threads || c.split != d->tile_split ||
                                        c.expand != d->tile_expand))
                continue;
            void* cnt = nullptr;
            RPP_TRY(workspace(kSlotCounter, 256, &cnt));
            return c.forward(B, *d, order, static_cast<unsigned*>(cnt), stream);
        }
        return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled chain tile matches");
    }
    if (d->tile_kind == 2) {
        for (const auto& c : thread_tiles<T>()) {
            if (c.width != d->basis.width || c.depth != d->basis.depth) continue;
            if (d->tile_threads > 0 && (c.threads != d->tile_threads || c.split != d->tile_split)) continue;
            return c.forward(B, *d, order, stream);
        }
        return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled thread-subtree tile matches");
    }
    if (const RegTile<T>* tile = pick_tile<T>(d->basis, d->tile_threads, d->tile_split, d->tile_expand, false)) {
        void* cnt = nullptr;
        RPP_TRY(workspace(kSlotCounter, 256, &cnt));
        return tile->forward(B, *d, order, static_cast<unsigned*>(cnt), stream);
    }
    if (d->tile_threads > 0) return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled tile matches the requested one");
    return signature_forward_generic<T>(B, *d, order, stream);
}

template <typename T>
int fmexp_adjoint_impl(const rpp_hip_ft_fmexp_adjoint_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    const int64_t n = d->num_batches;
    if (n < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_batches");
    if (n == 0) return 0;
    RPP_TRY(check_view(d->multiplier, d->basis, n, "multiplier", false));
    RPP_TRY(check_view(d->exponent, d->basis, n, "exponent", false));
    RPP_TRY(check_view(d->cotangent, d->basis, n, "cotangent", false));
    RPP_TRY(check_view(d->grad_multiplier, d->basis, n, "grad_multiplier", true));
    RPP_TRY(check_view(d->grad_exponent, d->basis, n, "grad_exponent", true));
    if (d->basis.depth < 1) return fail(RPP_HIP_INVALID_ARGUMENT, "ft_fmexp adjoint: depth must be >= 1");
    if (d->exponent.max_degree != 1 || d->exponent.min_degree > 1 || d->grad_exponent.max_degree != 1 ||
        d->grad_exponent.min_degree > 1)
        return fail(RPP_HIP_INVALID_ARGUMENT, "ft_fmexp adjoint: exponent windows must end at degree 1");
    if (d->grad_multiplier.min_degree != 0 || d->grad_multiplier.max_degree != d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "ft_fmexp adjoint: grad_multiplier must hold [0, depth]");
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    for (const rpp_hip_tensor_batch* o : {&d->grad_multiplier, &d->grad_exponent})
        for (const rpp_hip_tensor_batch* i : {&d->multiplier, &d->exponent, &d->cotangent})
            RPP_TRY(check_disjoint<T>(*o, *i, B.start, n, "ft_fmexp adjoint: a gradient and an input"));
    RPP_TRY(check_disjoint<T>(d->grad_multiplier, d->grad_exponent, B.start, n, "ft_fmexp adjoint: the two gradients"));
    hipStream_t stream = as_stream(s);
    const bool contig = d->multiplier.coeff_stride == 1 && d->cotangent.coeff_stride == 1 &&
                        d->grad_multiplier.coeff_stride == 1 && d->exponent.coeff_stride == 1;
    if (contig) {
        if (const RegTile<T>* tile = fmexp_adj_tile<T>(d->basis.width, d->basis.depth))
            return tile->fmexp_adjoint(B, *d, stream);
    }
    return fmexp_adjoint_generic<T>(B, *d, stream);
}

template <typename T>
int signature_backward_impl(const rpp_hip_signature_backward_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    if (d->num_paths < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_paths");
    if (d->num_paths == 0) return 0;
    RPP_TRY(check_view(d->signatures, d->basis, d->num_paths, "signatures", false));
    RPP_TRY(check_view(d->cotangent, d->basis, d->num_paths, "cotangent", false));
    if (d->signatures.min_degree != 0 || d->signatures.max_degree != d->basis.depth ||
        d->cotangent.min_degree != 0 || d->cotangent.max_degree != d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "signature backward: full windows [0, depth] required");
    if (!d->increments || !d->offsets || !d->grad_increments)
        return fail(RPP_HIP_INVALID_ARGUMENT, "null increments, offsets or gradient");
    if (d->checkpoint_every < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative checkpoint_every");
    if (d->checkpoint_every > 0 && (!d->checkpoints || !d->ckpt_offsets))
        return fail(RPP_HIP_INVALID_ARGUMENT, "checkpoints requested without buffers");
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    hipStream_t stream = as_stream(s);
    const int64_t* order = d->order;
    int64_t total = d->num_increments;
    if (order == nullptr) {
        std::vector<int64_t> lengths;
        int64_t* host_order = nullptr;
        // the tiled kernels index their partial sums by absolute row, so the
        // buffer spans rows [0, offsets[n]), not only the sum of the lengths
        RPP_TRY(path_order(d->offsets, d->num_paths, lengths, &host_order, stream, &total));
        order = host_order;
    } else if (total < 0) {
        return fail(RPP_HIP_INVALID_ARGUMENT, "num_increments must be >= offsets[num_paths] when order is given");
    }
    if (d->scratch == nullptr && d->tile_threads >= 0 && d->tile_kind == 2) {
        for (const auto& c : thread_tiles<T>()) {
            if (c.width != d->basis.width || c.depth != d->basis.depth || c.backward == nullptr) continue;
            if (d->tile_threads > 0 && (c.threads != d->tile_threads || c.split != d->tile_split)) continue;
            return c.backward(B, *d, order, total, stream, nullptr);
        }
        return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled thread-subtree backward tile matches");
    }
    if (d->scratch == nullptr && d->tile_threads >= 0) {
        if (const RegTile<T>* tile =
                pick_tile<T>(d->basis, d->tile_threads, d->tile_split, d->tile_expand, true)) {
            void* cnt = nullptr;
            RPP_TRY(workspace(kSlotCounter, 256, &cnt));
            return tile->backward(B, *d, order, total, static_cast<unsigned*>(cnt), stream);
        }
        if (d->tile_threads > 0) return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled backward tile matches");
    }
    return signature_backward_generic<T>(B, *d, order, stream);
}

// FP32 paths of at least this many increments get sqrt(L) checkpoints by
// default: measured on MI355X against the FP64 oracle, the exp(-x)
// reconstruction multiplies the FP32 gradient error by 1.2-3.6 from 32
// increments up (widths 3, 5, 6 at depth 6-8, 32-1024 increments), while in
// FP64 it stays below 5e-13 at 1024 increments.
constexpr int64_t kFp32CheckpointFrom = 32;

// The backward of signature_backward_impl with the thread-owned subtree
// tiles, allocating nothing: the per-row partial sums live in the fixed
// device scratch (noalloc_scratch), and the queue order must be the caller's.
template <typename T>
int signature_backward_noalloc_impl(const rpp_hip_signature_backward_desc* d, void* s) {
    if (!d) return fail(RPP_HIP_INVALID_ARGUMENT, "null descriptor");
    RPP_TRY(check_basis(d->basis));
    if (d->num_paths < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative num_paths");
    if (d->num_paths == 0) return 0;
    RPP_TRY(check_view(d->signatures, d->basis, d->num_paths, "signatures", false));
    RPP_TRY(check_view(d->cotangent, d->basis, d->num_paths, "cotangent", false));
    if (d->signatures.min_degree != 0 || d->signatures.max_degree != d->basis.depth ||
        d->cotangent.min_degree != 0 || d->cotangent.max_degree != d->basis.depth)
        return fail(RPP_HIP_INVALID_ARGUMENT, "signature backward: full windows [0, depth] required");
    if (!d->increments || !d->offsets || !d->grad_increments)
        return fail(RPP_HIP_INVALID_ARGUMENT, "null increments, offsets or gradient");
    if (d->checkpoint_every < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "negative checkpoint_every");
    if (d->checkpoint_every > 0 && (!d->checkpoints || !d->ckpt_offsets))
        return fail(RPP_HIP_INVALID_ARGUMENT, "checkpoints requested without buffers");
    if (!d->order)
        return fail(RPP_HIP_INVALID_ARGUMENT,
                    "signature backward without allocation: order (the device queue order) is required");
    if (d->num_increments < 0)
        return fail(RPP_HIP_INVALID_ARGUMENT, "num_increments must be >= offsets[num_paths] when order is given");
    if (d->scratch != nullptr)
        return fail(RPP_HIP_INVALID_ARGUMENT, "signature backward without allocation takes no scratch buffer");
    if (d->tile_threads < 0 || (d->tile_threads > 0 && d->tile_kind != 2))
        return fail(RPP_HIP_NOT_IMPLEMENTED,
                    "signature backward without allocation serves the thread-owned subtree tiles (tile_kind 2)");
    for (const auto& c : thread_tiles<T>()) {
        if (c.width != d->basis.width || c.depth != d->basis.depth || c.backward == nullptr) continue;
        if (d->tile_threads > 0 && (c.threads != d->tile_threads || c.split != d->tile_split)) continue;
        const size_t need = thread_backward_partial_bytes<T>(d->basis.width, d->num_increments);
        if (need > kNoAllocScratchBytes)
            return fail(RPP_HIP_INVALID_ARGUMENT,
                        "signature backward without allocation needs " + std::to_string(need) +
                            " bytes of partial sums (num_increments x width^2 values), above its fixed " +
                            std::to_string(kNoAllocScratchBytes) + "-byte scratch");
        void* ptr = nullptr;
        RPP_TRY(noalloc_scratch(&ptr));
        const DevBasis B = make_basis(d->basis.width, d->basis.depth);
        return c.backward(B, *d, d->order, d->num_increments, as_stream(s), ptr);
    }
    return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled thread-subtree backward tile matches");
}

}  // namespace
}  // namespace rpp_hip

extern "C" {

int32_t rpp_hip_signature_checkpoint_every(int32_t dtype_bytes, int64_t max_length) {
    if (dtype_bytes != 4 && dtype_bytes != 8)
        return -rpp_hip::fail(RPP_HIP_INVALID_ARGUMENT, "dtype_bytes must be 4 or 8");
    if (max_length < 0) return -rpp_hip::fail(RPP_HIP_INVALID_ARGUMENT, "negative max_length");
    if (max_length > (int64_t{1} << 60)) return -rpp_hip::fail(RPP_HIP_OUT_OF_BOUNDS, "max_length above 2^60");
    if (dtype_bytes == 8 || max_length < rpp_hip::kFp32CheckpointFrom) return 0;
    int64_t k = static_cast<int64_t>(std::sqrt(static_cast<double>(max_length)));
    while (k > 1 && (k - 1) * (k - 1) >= max_length) --k;
    while (k * k < max_length) ++k;
    return static_cast<int32_t>(k);
}

int rpp_hip_signature_backward_noalloc_f32(const rpp_hip_signature_backward_desc* d, void* s) {
    return signature_backward_noalloc_impl<float>(d, s);
}
int rpp_hip_signature_backward_noalloc_f64(const rpp_hip_signature_backward_desc* d, void* s) {
    return signature_backward_noalloc_impl<double>(d, s);
}

int rpp_hip_ft_fmexp_adjoint_f32(const rpp_hip_ft_fmexp_adjoint_desc* d, void* s) {
    return fmexp_adjoint_impl<float>(d, s);
}
int rpp_hip_ft_fmexp_adjoint_f64(const rpp_hip_ft_fmexp_adjoint_desc* d, void* s) {
    return fmexp_adjoint_impl<double>(d, s);
}
int rpp_hip_signature_f32(const rpp_hip_signature_desc* d, void* s) { return signature_impl<float>(d, s); }
int rpp_hip_signature_f64(const rpp_hip_signature_desc* d, void* s) { return signature_impl<double>(d, s); }
int rpp_hip_signature_backward_f32(const rpp_hip_signature_backward_desc* d, void* s) {
    return signature_backward_impl<float>(d, s);
}
int rpp_hip_signature_backward_f64(const rpp_hip_signature_backward_desc* d, void* s) {
    return signature_backward_impl<double>(d, s);
}

int64_t rpp_hip_signature_backward_scratch_bytes(const rpp_hip_signature_backward_desc* d, int32_t dtype_bytes) {
    if (!d || check_basis(d->basis)) return -1;
    const DevBasis B = make_basis(d->basis.width, d->basis.depth);
    const int64_t grid = std::min<int64_t>(std::max<int64_t>(d->num_paths, 1), device_cu_count() * 4);
    return backward_slot_elems(B) * grid * dtype_bytes + 256;
}

}  // extern "C"

extern "C" int rpp_hip_fmexp_tile(int32_t dtype_bytes, int32_t width, int32_t depth, int32_t* threads,
                                  int32_t* split) {
    auto report = [&](const auto* t) {
        if (!t) return 0;
        if (threads) *threads = t->threads;
        if (split) *split = t->split;
        return t->kind;
    };
    if (dtype_bytes == 4) return report(rpp_hip::fmexp_coal_tile<float>(width, depth));
    if (dtype_bytes == 8) return report(rpp_hip::fmexp_coal_tile<double>(width, depth));
    return rpp_hip::fail(RPP_HIP_INVALID_ARGUMENT, "dtype_bytes must be 4 or 8");
}

extern "C" int rpp_hip_signature_tiles(int32_t dtype_bytes, int32_t* out, int32_t capacity) {
    int n = 0;
    auto put = [&](int w, int d, int th, int sp, int kind, int ex) {
        if (n < capacity && out) {
            int32_t* o = out + 6 * n;
            o[0] = w, o[1] = d, o[2] = th, o[3] = sp, o[4] = kind, o[5] = ex;
        }
        ++n;
    };
    auto emit = [&](const auto& reg, const auto& chain, const auto& thread) {
        for (const auto& t : chain) put(t.width, t.depth, t.threads, t.split, 1, t.expand);
        for (const auto& t : reg) put(t.width, t.depth, t.threads, t.split, 0, t.reg_levels);
        for (const auto& t : thread) put(t.width, t.depth, t.threads, t.split, 2, t.backward ? 1 : 0);
    };
    if (dtype_bytes == 4)
        emit(rpp_hip::reg_tiles<float>(), rpp_hip::chain_tiles<float>(), rpp_hip::thread_tiles<float>());
    else if (dtype_bytes == 8)
        emit(rpp_hip::reg_tiles<double>(), rpp_hip::chain_tiles<double>(), rpp_hip::thread_tiles<double>());
    else return rpp_hip::fail(RPP_HIP_INVALID_ARGUMENT, "dtype_bytes must be 4 or 8");
    return n;
}
