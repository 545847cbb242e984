// Drop-in check: the reference's own free functions (rpp::ops::ft_mul, ...)
// called with rpp::hip::strategies::BlockStrategy on device buffers, against
// the same calls with rpp::cpu::strategies::SingleThreadStrategy on host
// copies of identical seeded inputs.  Built against the reference headers by
// oracle/Makefile (target dropin) into oracle/_ref/dropin_check; prints one
// line per operation and exits non-zero when any exceeds the tolerance
// |a - b| <= tol * max(1, |a|, |b|), tol = 1e-12 (double) / 1e-5 (float).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include <rpp/basis.hpp>
#include <rpp/basis/hall_basis.hpp>
#include <rpp/basis/lie_multiplication.hpp>
#include <rpp/cpu/single_thread/operations.hpp>
#include <rpp/cpu/strategies.hpp>
#include <rpp/operations.hpp>
#include <rpp/sparse/l2t_matrix.hpp>
#include <rpp/sparse/matrix.hpp>
#include <rpp/sparse/t2l_matrix.hpp>
#include <rpp/views/views.hpp>

#include <rpp/hip/block_strategy.hpp>

namespace {

using Arch = rpp::arch::NativeArchitecture;
using Index = Arch::Index;
using Degree = Arch::Degree;

int g_failures = 0;

template <typename T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    explicit Dev(const std::vector<T>& h) : n(h.size()) {
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) std::abort();
        if (n && hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) std::abort();
    }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    Dev(Dev&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    std::vector<T> host() const {
        std::vector<T> h(n);
        if (hipDeviceSynchronize() != hipSuccess) std::abort();
        if (n && hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) std::abort();
        return h;
    }
};

template <typename T>
void report(const char* what, const std::vector<T>& a, const std::vector<T>& b) {
    const double tol = std::is_same_v<T, double> ? 1e-12 : 1e-5;
    double worst = 0;
    bool nan_mismatch = false;
    for (size_t i = 0; i < a.size(); ++i) {
        const double x = a[i], y = b[i];
        if (std::isnan(x) || std::isnan(y)) {
            nan_mismatch |= std::isnan(x) != std::isnan(y);
            continue;
        }
        worst = std::max(worst, std::abs(x - y) / std::max({1.0, std::abs(x), std::abs(y)}));
    }
    const bool ok = worst <= tol && !nan_mismatch;
    g_failures += ok ? 0 : 1;
    std::printf("%-28s %-6s max_rel_err %.3e %s\n", what, std::is_same_v<T, double> ? "f64" : "f32", worst,
                ok ? "ok" : "FAIL");
}

template <typename T>
std::vector<T> random_batch(std::mt19937_64& rng, Index n, Index dim, double scale) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<T> v(static_cast<size_t>(n * dim));
    for (Index b = 0; b < n; ++b) {
        for (Index i = 0; i < dim; ++i) v[b * dim + i] = static_cast<T>(u(rng) * (i == 0 ? 1.0 : scale));
        v[b * dim] = T(1);
    }
    return v;
}

template <typename T>
void run(int width, int depth, Index n) {
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
    std::mt19937_64 rng(0xD0 + width * 16 + depth);
    auto lhs = random_batch<T>(rng, n, dim, 0.5), rhs = random_batch<T>(rng, n, dim, 0.5),
         add = random_batch<T>(rng, n, dim, 0.5);
    std::vector<T> incr(static_cast<size_t>(n * dim), T(0));
    std::normal_distribution<double> g(0.0, 0.3);
    for (Index b = 0; b < n; ++b)
        for (int l = 0; l < width; ++l) incr[b * dim + 1 + l] = static_cast<T>(g(rng));
    const rpp::cpu::strategies::SingleThreadStrategy<T> cpu;
    const rpp::hip::strategies::BlockStrategy<T> gpu;
    auto hb = [&](T* p, Degree mn, Degree mx) { return rpp::make_tensor_batch(p, dim, basis, mn, mx); };
    auto hcb = [&](const T* p, Degree mn, Degree mx) { return rpp::make_tensor_batch(p, dim, basis, mn, mx); };
    std::vector<T> nanv(static_cast<size_t>(n * dim), std::numeric_limits<T>::quiet_NaN());
    char name[64];
    auto tag = [&](const char* op) {
        std::snprintf(name, sizeof(name), "%s (%d,%d)", op, width, depth);
        return name;
    };

    {  // ft_mul, beta, windows
        std::vector<T> h = nanv;
        (void)rpp::ops::ft_mul(cpu, {}, hb(h.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 1, D), basis, n,
                               T(0.5));
        Dev<T> o(nanv), l(lhs), r(rhs);
        auto err = rpp::ops::ft_mul(gpu, {}, hb(o.p, 0, D), hcb(l.p, 0, D), hcb(r.p, 1, D), basis, n, T(0.5));
        if (!err) std::printf("ft_mul error: %s\n", err.message().data());
        report(tag("ft_mul"), o.host(), h);
    }
    {  // ft_mul on batch-interleaved batches (rpp::hip::layouts::InterleavedLayout)
        std::vector<T> h = nanv, li(lhs.size()), ri(rhs.size());
        for (Index bb = 0; bb < n; ++bb)
            for (Index i = 0; i < dim; ++i) {
                li[i * n + bb] = lhs[bb * dim + i];
                ri[i * n + bb] = rhs[bb * dim + i];
            }
        (void)rpp::ops::ft_mul(cpu, {}, hb(h.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 1, D), basis, n,
                               T(0.5));
        Dev<T> o(nanv), l(li), r(ri);
        auto il = [&](T* p, Degree mn, Degree mx) { return rpp::hip::make_interleaved_tensor_batch(p, n, basis, mn, mx); };
        auto err = rpp::ops::ft_mul(gpu, {}, il(o.p, 0, D), il(l.p, 0, D), il(r.p, 1, D), basis, n, T(0.5));
        if (!err) std::printf("ft_mul (interleaved) error: %s\n", err.message().data());
        std::vector<T> got = o.host(), back(got.size());
        for (Index bb = 0; bb < n; ++bb)
            for (Index i = 0; i < dim; ++i) back[bb * dim + i] = got[i * n + bb];
        report(tag("ft_mul (interleaved)"), back, h);
    }
    {  // ft_fma
        std::vector<T> h = nanv;
        (void)rpp::ops::ft_fma(cpu, {}, hb(h.data(), 0, D), hcb(add.data(), 0, 1), hcb(lhs.data(), 0, D),
                               hcb(rhs.data(), 0, D), basis, n, T(2), T(-1));
        Dev<T> o(nanv), a(add), l(lhs), r(rhs);
        (void)rpp::ops::ft_fma(gpu, {}, hb(o.p, 0, D), hcb(a.p, 0, 1), hcb(l.p, 0, D), hcb(r.p, 0, D), basis, n,
                               T(2), T(-1));
        report(tag("ft_fma"), o.host(), h);
    }
    using K = rpp::ops::FTInplaceFMAType;
    {
        std::vector<T> h = lhs;
        (void)rpp::ops::ft_inplace_fma<K::AEqualsBCPlusA>(cpu, {}, hb(h.data(), 0, D), hcb(rhs.data(), 0, D),
                                                          hcb(add.data(), 0, D), basis, n, T(0.5), T(1.5));
        Dev<T> a(lhs), b(rhs), c(add);
        (void)rpp::ops::ft_inplace_fma<K::AEqualsBCPlusA>(gpu, {}, hb(a.p, 0, D), hcb(b.p, 0, D), hcb(c.p, 0, D),
                                                          basis, n, T(0.5), T(1.5));
        report(tag("ft_inplace_fma<231>"), a.host(), h);
    }
    {
        std::vector<T> h = lhs;
        (void)rpp::ops::ft_inplace_fma<K::AEqualsABPlusC>(cpu, {}, hb(h.data(), 0, D), hcb(rhs.data(), 1, D),
                                                          hcb(add.data(), 0, D), basis, n, T(1), T(0.25));
        Dev<T> a(lhs), b(rhs), c(add);
        (void)rpp::ops::ft_inplace_fma<K::AEqualsABPlusC>(gpu, {}, hb(a.p, 0, D), hcb(b.p, 1, D), hcb(c.p, 0, D),
                                                          basis, n, T(1), T(0.25));
        report(tag("ft_inplace_fma<123>"), a.host(), h);
    }
    {
        std::vector<T> h = lhs;
        (void)rpp::ops::ft_inplace_fma<K::AEqualsBAPlusC>(cpu, {}, hb(h.data(), 0, D), hcb(rhs.data(), 0, D),
                                                          hcb(add.data(), 0, D), basis, n, T(1), T(1));
        Dev<T> a(lhs), b(rhs), c(add);
        (void)rpp::ops::ft_inplace_fma<K::AEqualsBAPlusC>(gpu, {}, hb(a.p, 0, D), hcb(b.p, 0, D), hcb(c.p, 0, D),
                                                          basis, n, T(1), T(1));
        report(tag("ft_inplace_fma<213>"), a.host(), h);
    }
    {
        std::vector<T> h = lhs;
        (void)rpp::ops::ft_inplace_mul(cpu, {}, hb(h.data(), 0, D), hcb(rhs.data(), 0, D), basis, n, T(3));
        Dev<T> a(lhs), b(rhs);
        (void)rpp::ops::ft_inplace_mul(gpu, {}, hb(a.p, 0, D), hcb(b.p, 0, D), basis, n, T(3));
        report(tag("ft_inplace_mul"), a.host(), h);
    }
    {  // exp, log of a full-window argument; fmexp with a degree-one exponent
        std::vector<T> he = nanv, hl = nanv, hf = nanv;
        (void)rpp::ops::ft_exp(cpu, {}, hb(he.data(), 0, D), hcb(rhs.data(), 0, D), basis, n);
        (void)rpp::ops::ft_log(cpu, {}, hb(hl.data(), 0, D), hcb(rhs.data(), 0, D), basis, n);
        (void)rpp::ops::ft_fmexp(cpu, {}, hb(hf.data(), 0, D), hcb(lhs.data(), 0, D), hcb(incr.data(), 1, 1), basis,
                                 n);
        Dev<T> oe(nanv), ol(nanv), of(nanv), r(rhs), l(lhs), x(incr);
        (void)rpp::ops::ft_exp(gpu, {}, hb(oe.p, 0, D), hcb(r.p, 0, D), basis, n);
        (void)rpp::ops::ft_log(gpu, {}, hb(ol.p, 0, D), hcb(r.p, 0, D), basis, n);
        (void)rpp::ops::ft_fmexp(gpu, {}, hb(of.p, 0, D), hcb(l.p, 0, D), hcb(x.p, 1, 1), basis, n);
        report(tag("ft_exp"), oe.host(), he);
        report(tag("ft_log"), ol.host(), hl);
        report(tag("ft_fmexp"), of.host(), hf);
    }
    {  // pairing into a scalar batch
        std::vector<T> h(static_cast<size_t>(n), T(0));
        (void)rpp::ops::tensor_pairing(cpu, {}, rpp::make_scalar_batch(h.data()), hcb(lhs.data(), 1, D),
                                       hcb(rhs.data(), 0, D), basis, n);
        Dev<T> o(h), l(lhs), r(rhs);
        (void)rpp::ops::tensor_pairing(gpu, {}, rpp::make_scalar_batch(o.p), hcb(l.p, 1, D), hcb(r.p, 0, D), basis,
                                       n);
        report(tag("tensor_pairing"), o.host(), h);
    }
    {  // adjoints
        std::vector<T> hl = nanv, hr = nanv;
        (void)rpp::ops::ft_adj_lmul(cpu, {}, hb(hl.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 0, D),
                                    basis, n);
        (void)rpp::ops::ft_adj_rmul(cpu, {}, hb(hr.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 0, D),
                                    basis, n);
        Dev<T> ol(nanv), orr(nanv), l(lhs), r(rhs);
        (void)rpp::ops::ft_adj_lmul(gpu, {}, hb(ol.p, 0, D), hcb(l.p, 0, D), hcb(r.p, 0, D), basis, n);
        (void)rpp::ops::ft_adj_rmul(gpu, {}, hb(orr.p, 0, D), hcb(l.p, 0, D), hcb(r.p, 0, D), basis, n);
        report(tag("ft_adj_lmul"), ol.host(), hl);
        report(tag("ft_adj_rmul"), orr.host(), hr);
    }
    {  // antipode and reflection, output window from degree 1
        std::vector<T> ha = nanv, hr = nanv;
        (void)rpp::ops::tensor_antipode(cpu, {}, hb(ha.data(), 1, D), hcb(rhs.data(), 0, D), basis, n);
        (void)rpp::ops::tensor_reflect(cpu, {}, hb(hr.data(), 1, D), hcb(rhs.data(), 0, D), basis, n);
        Dev<T> oa(nanv), orf(nanv), r(rhs);
        (void)rpp::ops::tensor_antipode(gpu, {}, hb(oa.p, 1, D), hcb(r.p, 0, D), basis, n);
        (void)rpp::ops::tensor_reflect(gpu, {}, hb(orf.p, 1, D), hcb(r.p, 0, D), basis, n);
        report(tag("tensor_antipode"), oa.host(), ha);
        report(tag("tensor_reflect"), orf.host(), hr);
    }
    {  // linear operations on partial windows
        const Degree lo = depth >= 2 ? 1 : 0;
        std::vector<T> hs = lhs, ha = nanv, hadd = nanv, hi = lhs, hc = lhs, hid = lhs, hai = lhs;
        (void)rpp::ops::vector_scalar_multiply(cpu, {}, hb(hs.data(), lo, D), basis, n, T(-0.5));
        (void)rpp::ops::vector_assign(cpu, {}, hb(ha.data(), 0, D), hcb(rhs.data(), lo, D), basis, n);
        (void)rpp::ops::vector_add(cpu, {}, hb(hadd.data(), 0, D), hcb(lhs.data(), lo, D), hcb(rhs.data(), 0, lo),
                                   basis, n, T(0.5), T(2));
        (void)rpp::ops::vector_inplace_add(cpu, {}, hb(hi.data(), 0, D), hcb(rhs.data(), lo, D), basis, n, T(-1));
        (void)rpp::ops::vector_set_constant(cpu, {}, hb(hc.data(), lo, D), basis, n, T(0.25));
        (void)rpp::ops::tensor_set_identity(cpu, {}, hb(hid.data(), 0, D), basis, n, T(3));
        (void)rpp::ops::tensor_add_identity(cpu, {}, hb(hai.data(), 0, D), basis, n, T(-1));
        Dev<T> s(lhs), a(nanv), ad(nanv), ip(lhs), c(lhs), id(lhs), ai(lhs), l(lhs), r(rhs);
        (void)rpp::ops::vector_scalar_multiply(gpu, {}, hb(s.p, lo, D), basis, n, T(-0.5));
        (void)rpp::ops::vector_assign(gpu, {}, hb(a.p, 0, D), hcb(r.p, lo, D), basis, n);
        (void)rpp::ops::vector_add(gpu, {}, hb(ad.p, 0, D), hcb(l.p, lo, D), hcb(r.p, 0, lo), basis, n, T(0.5),
                                   T(2));
        (void)rpp::ops::vector_inplace_add(gpu, {}, hb(ip.p, 0, D), hcb(r.p, lo, D), basis, n, T(-1));
        (void)rpp::ops::vector_set_constant(gpu, {}, hb(c.p, lo, D), basis, n, T(0.25));
        (void)rpp::ops::tensor_set_identity(gpu, {}, hb(id.p, 0, D), basis, n, T(3));
        (void)rpp::ops::tensor_add_identity(gpu, {}, hb(ai.p, 0, D), basis, n, T(-1));
        report(tag("vector_scalar_multiply"), s.host(), hs);
        report(tag("vector_assign"), a.host(), ha);
        report(tag("vector_add"), ad.host(), hadd);
        report(tag("vector_inplace_add"), ip.host(), hi);
        report(tag("vector_set_constant"), c.host(), hc);
        report(tag("tensor_set_identity"), id.host(), hid);
        report(tag("tensor_add_identity"), ai.host(), hai);
    }
    if (depth <= 5) {  // the shuffle family; the CPU half-shuffle adjoints through their operation objects
        const Degree lo = depth >= 2 ? 1 : 0;
        std::vector<T> hm = nanv, hf = nanv, hi = add, ha = nanv, hh = nanv, hl = nanv, hr = nanv;
        (void)rpp::ops::st_mul(cpu, {}, hb(hm.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), lo, D), basis, n,
                               T(0.5));
        (void)rpp::ops::st_fma(cpu, {}, hb(hf.data(), 0, D), hcb(add.data(), lo, D), hcb(lhs.data(), 0, D),
                               hcb(rhs.data(), 0, D), basis, n, T(2), T(-1));
        (void)rpp::ops::st_inplace_fma(cpu, {}, hb(hi.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 0, D),
                                       basis, n, T(0.25), T(1));
        (void)rpp::ops::st_adj_mul(cpu, {}, hb(ha.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 0, D), basis,
                                   n);
        (void)rpp::ops::left_hs_mul(cpu, {}, hb(hh.data(), 0, D), hcb(lhs.data(), 0, D), hcb(rhs.data(), 0, D), basis,
                                    n, T(1));
        {
            using Strategy = rpp::cpu::strategies::SingleThreadStrategy<T>;
            auto const ctx = Strategy::make_context(nullptr);
            auto const pack = rpp::basis::make_basis_pack(basis);
            for (Index i = 0; i < n; ++i) {
                rpp::DenseTensorView<T*, rpp::basis::TensorBasis<Arch>> lo_v(hl.data() + i * dim, basis, 0, D);
                rpp::DenseTensorView<T*, rpp::basis::TensorBasis<Arch>> ro_v(hr.data() + i * dim, basis, 0, D);
                rpp::DenseTensorView<const T*, rpp::basis::TensorBasis<Arch>> x_v(lhs.data() + i * dim, basis, 0, D);
                rpp::DenseTensorView<const T*, rpp::basis::TensorBasis<Arch>> y_v(rhs.data() + i * dim, basis, 0, D);
                rpp::ops::LeftHSAdjLMul<Strategy>{ctx, pack}(ctx, lo_v, x_v, y_v);
                rpp::ops::LeftHSAdjRMul<Strategy>{ctx, pack}(ctx, ro_v, x_v, y_v);
            }
        }
        Dev<T> m(nanv), f(nanv), in(add), a(nanv), h(nanv), l(nanv), r(nanv), x(lhs), y(rhs), z(add);
        (void)rpp::ops::st_mul(gpu, {}, hb(m.p, 0, D), hcb(x.p, 0, D), hcb(y.p, lo, D), basis, n, T(0.5));
        (void)rpp::ops::st_fma(gpu, {}, hb(f.p, 0, D), hcb(z.p, lo, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n, T(2),
                               T(-1));
        (void)rpp::ops::st_inplace_fma(gpu, {}, hb(in.p, 0, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n, T(0.25),
                                       T(1));
        (void)rpp::ops::st_adj_mul(gpu, {}, hb(a.p, 0, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n);
        (void)rpp::ops::left_hs_mul(gpu, {}, hb(h.p, 0, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n, T(1));
        (void)rpp::ops::left_hs_adj_lmul(gpu, {}, hb(l.p, 0, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n);
        (void)rpp::ops::left_hs_adj_rmul(gpu, {}, hb(r.p, 0, D), hcb(x.p, 0, D), hcb(y.p, 0, D), basis, n);
        report(tag("st_mul"), m.host(), hm);
        report(tag("st_fma"), f.host(), hf);
        report(tag("st_inplace_fma"), in.host(), hi);
        report(tag("st_adj_mul"), a.host(), ha);
        report(tag("left_hs_mul"), h.host(), hh);
        report(tag("left_hs_adj_lmul"), l.host(), hl);
        report(tag("left_hs_adj_rmul"), r.host(), hr);
    }
    if (width >= 2) {  // linear operations on a Lie grading (the Hall-set LieBasis)
        rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width), D);
        auto lb = hall.to_lie_basis();
        const Index ls = lb.true_size();
        std::vector<T> x(static_cast<size_t>(n * ls)), y(static_cast<size_t>(n * ls));
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        for (auto& v : x) v = static_cast<T>(u(rng));
        for (auto& v : y) v = static_cast<T>(u(rng));
        auto lie_b = [&](T* p, Degree mn) { return rpp::make_lie_batch(p, ls, lb, mn, D); };
        auto lie_cb = [&](const T* p, Degree mn) { return rpp::make_lie_batch(p, ls, lb, mn, D); };
        std::vector<T> hadd(static_cast<size_t>(n * ls), T(7)), hasg = hadd;
        (void)rpp::ops::vector_add(cpu, {}, lie_b(hadd.data(), 1), lie_cb(x.data(), 1), lie_cb(y.data(), 2), lb, n,
                                   T(1), T(-3));
        (void)rpp::ops::vector_assign(cpu, {}, lie_b(hasg.data(), 1), lie_cb(y.data(), 2), lb, n);
        Dev<T> gadd(std::vector<T>(hadd.size(), T(7))), gasg(std::vector<T>(hadd.size(), T(7))), gx(x), gy(y);
        (void)rpp::ops::vector_add(gpu, {}, lie_b(gadd.p, 1), lie_cb(gx.p, 1), lie_cb(gy.p, 2), lb, n, T(1), T(-3));
        (void)rpp::ops::vector_assign(gpu, {}, lie_b(gasg.p, 1), lie_cb(gy.p, 2), lb, n);
        report(tag("vector_add (Lie)"), gadd.host(), hadd);
        report(tag("vector_assign (Lie)"), gasg.host(), hasg);
    }
    if (width >= 2) {  // tensor_to_lie / lie_to_tensor with the reference-built matrices
        rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width), D);
        auto lb = hall.to_lie_basis();
        rpp::LieMultiplicationCache<Arch> cache(lb);
        auto t2l = rpp::sparse::make_t2l_matrix<T>(lb, cache);
        auto t2l_csr = rpp::sparse::swap_format(t2l.view());
        auto l2t = rpp::sparse::make_l2t_matrix<T>(lb);
        const Index ls = lb.true_size();
        std::vector<T> hlie(static_cast<size_t>(n * ls), T(0)), hten = nanv;
        auto lie_b = [&](T* p) { return rpp::make_lie_batch(p, ls, lb, Degree{1}, D); };
        auto lie_cb = [&](const T* p) { return rpp::make_lie_batch(p, ls, lb, Degree{1}, D); };
        (void)rpp::ops::tensor_to_lie(cpu, {}, lie_b(hlie.data()), hcb(rhs.data(), 0, D), lb, basis, n,
                                      t2l_csr.view());
        (void)rpp::ops::lie_to_tensor(cpu, {}, hb(hten.data(), 0, D), lie_cb(hlie.data()), basis, lb, n,
                                      l2t.view());
        auto upload = [](auto const& m) {
            using V = std::remove_cv_t<std::remove_reference_t<decltype(*m.data())>>;
            std::vector<V> v(m.data(), m.data() + m.nnz());
            std::vector<Index> i(m.indices(), m.indices() + m.nnz());
            std::vector<Index> o(m.offsets(), m.offsets() + m.outer_dim() + 1);
            return std::make_tuple(Dev<V>(v), Dev<Index>(i), Dev<Index>(o));
        };
        auto [tv, ti, to] = upload(t2l_csr);
        auto [lv, li, lo] = upload(l2t);
        auto d_t2l = rpp::sparse::make_csr_matrix(static_cast<const T*>(tv.p), static_cast<const Index*>(ti.p),
                                                  static_cast<const Index*>(to.p), t2l_csr.nnz(),
                                                  t2l_csr.outer_dim(), t2l_csr.inner_dim());
        auto d_l2t = rpp::sparse::make_csc_matrix(static_cast<const T*>(lv.p), static_cast<const Index*>(li.p),
                                                  static_cast<const Index*>(lo.p), l2t.nnz(), l2t.inner_dim(),
                                                  l2t.outer_dim());
        Dev<T> glie(std::vector<T>(static_cast<size_t>(n * ls), T(0))), gten(nanv), r(rhs);
        (void)rpp::ops::tensor_to_lie(gpu, {}, lie_b(glie.p), hcb(r.p, 0, D), lb, basis, n,
                                      rpp::sparse::GradedMatrixView<rpp::sparse::MatrixFormat::CSR, const T*,
                                                                    const Index*, const Index*>(
                                          d_t2l.data(), d_t2l.indices(), d_t2l.offsets(), d_t2l.nnz(),
                                          d_t2l.outer_dim(), d_t2l.inner_dim()));
        (void)rpp::ops::lie_to_tensor(gpu, {}, hb(gten.p, 0, D), lie_cb(glie.p), basis, lb, n,
                                      rpp::sparse::GradedMatrixView<rpp::sparse::MatrixFormat::CSC, const T*,
                                                                    const Index*, const Index*>(
                                          d_l2t.data(), d_l2t.indices(), d_l2t.offsets(), d_l2t.nnz(),
                                          d_l2t.outer_dim(), d_l2t.inner_dim()));
        report(tag("tensor_to_lie (CSR)"), glie.host(), hlie);
        report(tag("lie_to_tensor (CSC)"), gten.host(), hten);
        // the CSC T2L exactly as make_t2l_matrix returns it, through
        // tensor_to_lie and through sparse_matrix_vector_product with alpha
        auto [cv, ci, co] = upload(t2l);
        rpp::sparse::GradedMatrixView<rpp::sparse::MatrixFormat::CSC, const T*, const Index*, const Index*> d_t2l_csc(
            static_cast<const T*>(cv.p), static_cast<const Index*>(ci.p), static_cast<const Index*>(co.p), t2l.nnz(),
            t2l.outer_dim(), t2l.inner_dim());
        std::vector<T> hlie2(static_cast<size_t>(n * ls), T(0)), hlie3 = hlie2;
        (void)rpp::ops::tensor_to_lie(cpu, {}, lie_b(hlie2.data()), hcb(rhs.data(), 0, D), lb, basis, n, t2l.view());
        (void)rpp::ops::sparse_matrix_vector_product(cpu, {}, lie_b(hlie3.data()), hcb(rhs.data(), 0, D), lb, basis,
                                                     n, t2l.view(), T(0.5));
        Dev<T> glie2(std::vector<T>(static_cast<size_t>(n * ls), T(0))), glie3(glie2.host());
        (void)rpp::ops::tensor_to_lie(gpu, {}, lie_b(glie2.p), hcb(r.p, 0, D), lb, basis, n, d_t2l_csc);
        (void)rpp::ops::sparse_matrix_vector_product(gpu, {}, lie_b(glie3.p), hcb(r.p, 0, D), lb, basis, n,
                                                     d_t2l_csc, T(0.5));
        report(tag("tensor_to_lie (CSC)"), glie2.host(), hlie2);
        report(tag("sparse_matrix_vector_product (CSC, alpha 0.5)"), glie3.host(), hlie3);
        // the library caches what it derived from a matrix by its device
        // arrays: forget them before they are freed
        for (const void* v : {static_cast<const void*>(tv.p), static_cast<const void*>(lv.p),
                              static_cast<const void*>(cv.p)})
            (void)rpp_hip_sparse_matrix_forget(v);
    }
}

}  // namespace

int main() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        std::printf("no HIP device\n");
        return 2;
    }
    for (auto [w, d] : {std::pair{2, 5}, std::pair{3, 4}, std::pair{4, 4}, std::pair{5, 3}, std::pair{1, 6}}) {
        run<double>(w, d, 7);
        run<float>(w, d, 7);
    }
    std::printf("%s\n", g_failures ? "DROPIN FAIL" : "DROPIN OK");
    return g_failures ? 1 : 0;
}
