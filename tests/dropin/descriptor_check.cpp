// Host-only check of the hip::block drop-in layer (include/rpp/hip/
// block_strategy.hpp): the reference's own free functions are called with
// hip::strategies::BlockStrategy on object-major (StrideLayout) and
// batch-interleaved (rpp::hip::layouts::InterleavedLayout) batches, and every
// descriptor the layer hands to the C-ABI is checked field by field.
//
// The C-ABI entry points the calls below reach are defined here as recorders:
// each stores a copy of its descriptor and returns the code the check asks
// for, so nothing is launched and no GPU is needed.  This binary is built
// from the reference's headers and runs only where they are; it is test
// infrastructure, never linked with librpp_hip.so.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <rpp/basis/tensor_basis.hpp>
#include <rpp/basis/hall_basis.hpp>
#include <rpp/operations.hpp>
#include <rpp/sparse/matrix.hpp>
#include <rpp/views/views.hpp>

#include <rpp/hip/block_strategy.hpp>

namespace {

std::string g_name;
std::vector<unsigned char> g_desc;
void* g_stream = nullptr;
int g_return = 0;
const char* g_message = "";
int g_failures = 0;

int record(const char* name, const void* desc, size_t size, void* stream) {
    g_name = name;
    g_desc.assign(static_cast<const unsigned char*>(desc), static_cast<const unsigned char*>(desc) + size);
    g_stream = stream;
    return g_return;
}

template <typename D>
D last(const char* name) {
    D d{};
    if (g_name != name) {
        std::printf("FAIL expected a call of %s, got %s\n", name, g_name.c_str());
        ++g_failures;
        return d;
    }
    std::memcpy(&d, g_desc.data(), sizeof(D));
    return d;
}

void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++g_failures;
    }
}

void check_tensor(const rpp_hip_tensor_batch& t, const void* data, int64_t bs, int64_t cs, int mn, int mx,
                  const char* what) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "%s: data %p/%p strides %lld,%lld/%lld,%lld window %d,%d/%d,%d", what, t.data,
                  data, static_cast<long long>(t.batch_stride), static_cast<long long>(bs),
                  static_cast<long long>(t.coeff_stride), static_cast<long long>(cs), t.min_degree, mn,
                  t.max_degree, mx);
    check(t.data == data && t.batch_stride == bs && t.coeff_stride == cs && t.min_degree == mn && t.max_degree == mx,
          msg);
}

}  // namespace

#define RPP_RECORD(NAME, DESC)                                                                   \
    extern "C" int NAME##_f32(const DESC* d, void* s) { return record(#NAME "_f32", d, sizeof(DESC), s); } \
    extern "C" int NAME##_f64(const DESC* d, void* s) { return record(#NAME "_f64", d, sizeof(DESC), s); }

RPP_RECORD(rpp_hip_ft_mul, rpp_hip_ft_mul_desc)
RPP_RECORD(rpp_hip_ft_fmexp, rpp_hip_ft_fmexp_desc)
RPP_RECORD(rpp_hip_ft_log, rpp_hip_ft_unary_desc)
RPP_RECORD(rpp_hip_tensor_pairing, rpp_hip_tensor_pairing_desc)
RPP_RECORD(rpp_hip_sparse_matrix_vector_product, rpp_hip_spmv_desc)
RPP_RECORD(rpp_hip_tensor_to_lie_matrix, rpp_hip_lie_matrix_convert_desc)
RPP_RECORD(rpp_hip_lie_to_tensor_matrix, rpp_hip_lie_matrix_convert_desc)

extern "C" const char* rpp_hip_last_error(void) { return g_message; }

namespace {

using Arch = rpp::arch::NativeArchitecture;
using Index = Arch::Index;
using Degree = Arch::Degree;

template <typename T>
void run(int width, int depth) {
    const char* sfx = std::is_same_v<T, double> ? "_f64" : "_f32";
    auto name = [&](const char* op) { return std::string(op) + sfx; };
    std::vector<Index> db(static_cast<size_t>(depth) + 2);
    Index pw = 1;
    db[0] = 0;
    for (int k = 1; k <= depth + 1; ++k) {
        db[k] = db[k - 1] + pw;
        pw *= width;
    }
    rpp::basis::TensorBasis<Arch> basis(static_cast<Degree>(width), static_cast<Degree>(depth), db.data());
    const Index dim = db[depth + 1];
    const auto D = static_cast<Degree>(depth);
    const Index n = 7;
    std::vector<T> a(static_cast<size_t>(n * dim)), b(a.size()), c(a.size());
    const rpp::hip::strategies::BlockStrategy<T> hip;
    void* stream = reinterpret_cast<void*>(0x5a5a);
    auto om = [&](T* p, Degree mn, Degree mx) { return rpp::make_tensor_batch(p, dim, basis, mn, mx); };
    auto il = [&](T* p, Degree mn, Degree mx) {
        return rpp::hip::make_interleaved_tensor_batch(p, n, basis, mn, mx);
    };

    // ft_mul: object-major out, interleaved lhs, object-major rhs, beta
    (void)rpp::ops::ft_mul(hip, {stream}, om(a.data(), 1, D), il(b.data(), 0, D - 1), om(c.data(), 1, D), basis, n,
                           T(0.5));
    {
        auto d = last<rpp_hip_ft_mul_desc>(name("rpp_hip_ft_mul").c_str());
        check(d.basis.width == width && d.basis.depth == depth && d.num_batches == n && d.beta == 0.5,
              "ft_mul basis, count, beta");
        check_tensor(d.out, a.data(), dim, 1, 1, depth, "ft_mul out (object-major)");
        check_tensor(d.lhs, b.data(), 1, n, 0, depth - 1, "ft_mul lhs (interleaved)");
        check_tensor(d.rhs, c.data(), dim, 1, 1, depth, "ft_mul rhs (object-major)");
        check(g_stream == stream, "ft_mul stream");
    }
    // ft_fmexp: interleaved out and multiplier, degree-one exponent view
    (void)rpp::ops::ft_fmexp(hip, {stream}, il(a.data(), 0, D), il(b.data(), 0, D), om(c.data(), 1, 1), basis, n);
    {
        auto d = last<rpp_hip_ft_fmexp_desc>(name("rpp_hip_ft_fmexp").c_str());
        check_tensor(d.out, a.data(), 1, n, 0, depth, "ft_fmexp out (interleaved)");
        check_tensor(d.multiplier, b.data(), 1, n, 0, depth, "ft_fmexp multiplier (interleaved)");
        check_tensor(d.exponent, c.data(), dim, 1, 1, 1, "ft_fmexp exponent");
    }
    // ft_log, interleaved both sides
    (void)rpp::ops::ft_log(hip, {stream}, il(a.data(), 0, D), il(b.data(), 0, D), basis, n);
    {
        auto d = last<rpp_hip_ft_unary_desc>(name("rpp_hip_ft_log").c_str());
        check_tensor(d.out, a.data(), 1, n, 0, depth, "ft_log out");
        check_tensor(d.arg, b.data(), 1, n, 0, depth, "ft_log arg");
    }
    // tensor_pairing: scalar batch out, interleaved argument
    std::vector<T> s(static_cast<size_t>(n));
    (void)rpp::ops::tensor_pairing(hip, {stream}, rpp::make_scalar_batch(s.data()), om(c.data(), 0, D),
                                   il(b.data(), 1, D), basis, n);
    {
        auto d = last<rpp_hip_tensor_pairing_desc>(name("rpp_hip_tensor_pairing").c_str());
        check(d.out.data == s.data() && d.out.stride == 1, "tensor_pairing out");
        check_tensor(d.functional, c.data(), dim, 1, 0, depth, "tensor_pairing functional");
        check_tensor(d.arg, b.data(), 1, n, 1, depth, "tensor_pairing arg (interleaved)");
    }
    // tensor_to_lie / lie_to_tensor / sparse_matrix_vector_product with a CSC
    // matrix (arrays never read here) between an interleaved tensor batch and
    // an object-major Lie batch
    if (width >= 2) {
        rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width), D);
        auto lb = hall.to_lie_basis();
        const Index ls = lb.true_size();
        std::vector<T> lie(static_cast<size_t>(n * ls)), vals(5);
        std::vector<Index> idx(5), offs(static_cast<size_t>(dim) + 1);
        rpp::sparse::GradedMatrixView<rpp::sparse::MatrixFormat::CSC, const T*, const Index*, const Index*> m(
            vals.data(), idx.data(), offs.data(), 5, dim, ls - 1);
        auto lie_b = [&](T* p) { return rpp::make_lie_batch(p, ls, lb, Degree{1}, D); };
        (void)rpp::ops::tensor_to_lie(hip, {stream}, lie_b(lie.data()), il(b.data(), 0, D), lb, basis, n, m);
        {
            auto d = last<rpp_hip_lie_matrix_convert_desc>(name("rpp_hip_tensor_to_lie_matrix").c_str());
            check(d.basis.width == width && d.basis.depth == depth && d.num_batches == n, "tensor_to_lie basis");
            check_tensor(d.tensor, b.data(), 1, n, 0, depth, "tensor_to_lie tensor (interleaved)");
            check_tensor(d.lie, lie.data(), ls, 1, 1, depth, "tensor_to_lie lie");
            check(d.matrix.format == RPP_HIP_CSC && d.matrix.values == vals.data() && d.matrix.indices == idx.data() &&
                      d.matrix.offsets == offs.data() && d.matrix.nnz == 5 && d.matrix.rows == m.rows() &&
                      d.matrix.cols == m.cols(),
                  "tensor_to_lie matrix");
        }
        (void)rpp::ops::lie_to_tensor(hip, {stream}, il(a.data(), 0, D), lie_b(lie.data()), basis, lb, n, m);
        {
            auto d = last<rpp_hip_lie_matrix_convert_desc>(name("rpp_hip_lie_to_tensor_matrix").c_str());
            check_tensor(d.tensor, a.data(), 1, n, 0, depth, "lie_to_tensor tensor (interleaved)");
            check_tensor(d.lie, lie.data(), ls, 1, 1, depth, "lie_to_tensor lie");
        }
        (void)rpp::ops::sparse_matrix_vector_product(hip, {stream}, lie_b(lie.data()), il(b.data(), 2, D), lb, basis,
                                                     n, m, T(-2));
        {
            auto d = last<rpp_hip_spmv_desc>(name("rpp_hip_sparse_matrix_vector_product").c_str());
            check(d.num_batches == n && d.alpha == -2.0 && d.matrix.format == RPP_HIP_CSC, "spmv count, alpha");
            check(d.out.data == lie.data() && d.out.batch_stride == ls && d.out.coeff_stride == 1 &&
                      d.out.begin == lb.start_of_degree(1) && d.out.size == lb.end_of_degree(D) - lb.start_of_degree(1),
                  "spmv out (Lie view from degree 1)");
            check(d.arg.data == b.data() && d.arg.batch_stride == 1 && d.arg.coeff_stride == n &&
                      d.arg.begin == db[2] && d.arg.size == db[depth + 1] - db[2],
                  "spmv arg (interleaved tensor view from degree 2)");
        }
    }
    // a failing call maps the C-ABI code and message to rpp::Error
    g_return = RPP_HIP_INVALID_ARGUMENT;
    g_message = "recorded failure";
    auto err = rpp::ops::ft_log(hip, {stream}, il(a.data(), 0, D), il(b.data(), 0, D), basis, n);
    check(!err && err.code() == rpp::ErrorCode::InvalidArgument && std::string(err.message().data()) == g_message,
          "error mapping");
    g_return = 0;
    g_message = "";
}

}  // namespace

int main() {
    for (auto [w, d] : {std::pair{2, 5}, std::pair{3, 4}, std::pair{5, 3}}) {
        run<double>(w, d);
        run<float>(w, d);
    }
    std::printf("%s (%d failures)\n", g_failures ? "DESCRIPTORS FAIL" : "DESCRIPTORS OK", g_failures);
    return g_failures ? 1 : 0;
}
