// ref_driver.cpp -- C-ABI driver around the REAL reference strategy
// rpp::cpu::strategies::SingleThreadStrategy (RoughPathPrimitives v1.0.2).
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference's own headers where they lie (-I <reference>/include) into
// oracle/_ref/librpp_ref.so, which the tests use to pin the plain-C oracle and
// which bench.py times as the reference CPU baseline.  The reference headers
// are not copied anywhere; this file only calls the reference's public
// free functions (operations/**.hpp) through its batch/view model.
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
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

namespace {

using Arch = rpp::arch::NativeArchitecture;
using Index = Arch::Index;
using Degree = Arch::Degree;
using TBasis = rpp::basis::TensorBasis<Arch>;

struct BasisHolder {
    std::vector<Index> degree_begin;
    TBasis basis;
    static std::vector<Index> make(int width, int depth) {
        std::vector<Index> db(static_cast<size_t>(depth) + 2);
        db[0] = 0;
        Index pw = 1;
        for (int k = 1; k <= depth + 1; ++k) {
            db[k] = db[k - 1] + pw;
            pw *= width;
        }
        return db;
    }
    BasisHolder(int width, int depth)
        : degree_begin(make(width, depth)),
          basis(static_cast<Degree>(width), static_cast<Degree>(depth),
                degree_begin.data()) {}
};

template <typename T>
using Strat = rpp::cpu::strategies::SingleThreadStrategy<T>;

template <typename T, typename P>
auto batch(P* data, long long stride, TBasis const& basis, int mn, int mx) {
    return rpp::make_tensor_batch(data, static_cast<Index>(stride), basis,
                                  static_cast<Degree>(mn),
                                  static_cast<Degree>(mx));
}

// rpp::Error converts to true on success (support/error.hpp:140-142).
int code_of(rpp::Error<std::string> const& err) {
    return err ? 0 : static_cast<int>(err.code());
}

// Split [0,n) over `threads` std::threads, each calling fn(first, count).
template <typename Fn>
int split(long long n, int threads, Fn&& fn) {
    if (threads <= 1 || n < threads) return fn(0LL, n);
    std::vector<std::thread> pool;
    std::vector<int> codes(static_cast<size_t>(threads), 0);
    const long long per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const long long first = per * t;
        const long long count = first >= n ? 0 : std::min(per, n - first);
        pool.emplace_back([&, t, first, count] {
            if (count > 0) codes[static_cast<size_t>(t)] = fn(first, count);
        });
    }
    for (auto& th : pool) th.join();
    for (int c : codes)
        if (c) return c;
    return 0;
}

template <typename T>
int ft_mul(int w, int N, long long n, T* out, long long os, int omin, int omax,
           const T* lhs, long long ls, int lmin, int lmax, const T* rhs,
           long long rs, int rmin, int rmax, T beta, int threads) {
    BasisHolder bh(w, N);
    return split(n, threads, [&](long long f, long long c) {
        return code_of(rpp::ops::ft_mul(
            Strat<T>{}, {}, batch<T>(out + f * os, os, bh.basis, omin, omax),
            batch<T>(lhs + f * ls, ls, bh.basis, lmin, lmax),
            batch<T>(rhs + f * rs, rs, bh.basis, rmin, rmax), bh.basis,
            static_cast<Index>(c), beta));
    });
}

template <typename T>
int ft_fma(int w, int N, long long n, T* out, long long os, int omin, int omax,
           const T* a, long long as, int amin, int amax, const T* b,
           long long bs, int bmin, int bmax, const T* c, long long cs,
           int cmin, int cmax, T alpha, T beta) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::ft_fma(
        Strat<T>{}, {}, batch<T>(out, os, bh.basis, omin, omax),
        batch<T>(a, as, bh.basis, amin, amax),
        batch<T>(b, bs, bh.basis, bmin, bmax),
        batch<T>(c, cs, bh.basis, cmin, cmax), bh.basis,
        static_cast<Index>(n), alpha, beta));
}

template <typename T>
int ft_inplace_fma(int variant, int w, int N, long long n, T* a, long long as,
                   int amin, int amax, const T* b, long long bs, int bmin,
                   int bmax, const T* c, long long cs, int cmin, int cmax,
                   T alpha, T beta) {
    using K = rpp::ops::FTInplaceFMAType;
    BasisHolder bh(w, N);
    auto A = batch<T>(a, as, bh.basis, amin, amax);
    auto Bb = batch<T>(b, bs, bh.basis, bmin, bmax);
    auto C = batch<T>(c, cs, bh.basis, cmin, cmax);
    const auto nn = static_cast<Index>(n);
    if (variant == 231)
        return code_of(rpp::ops::ft_inplace_fma<K::AEqualsBCPlusA>(
            Strat<T>{}, {}, A, Bb, C, bh.basis, nn, alpha, beta));
    if (variant == 123)
        return code_of(rpp::ops::ft_inplace_fma<K::AEqualsABPlusC>(
            Strat<T>{}, {}, A, Bb, C, bh.basis, nn, alpha, beta));
    if (variant == 213)
        return code_of(rpp::ops::ft_inplace_fma<K::AEqualsBAPlusC>(
            Strat<T>{}, {}, A, Bb, C, bh.basis, nn, alpha, beta));
    return 3;
}

template <typename T>
int ft_inplace_mul(int w, int N, long long n, T* lhs, long long ls, int lmin,
                   int lmax, const T* rhs, long long rs, int rmin, int rmax,
                   T beta) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::ft_inplace_mul(
        Strat<T>{}, {}, batch<T>(lhs, ls, bh.basis, lmin, lmax),
        batch<T>(rhs, rs, bh.basis, rmin, rmax), bh.basis,
        static_cast<Index>(n), beta));
}

template <typename T>
int ft_exp(int w, int N, long long n, T* out, long long os, int omin, int omax,
           const T* arg, long long as, int amin, int amax) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::ft_exp(
        Strat<T>{}, {}, batch<T>(out, os, bh.basis, omin, omax),
        batch<T>(arg, as, bh.basis, amin, amax), bh.basis,
        static_cast<Index>(n)));
}

template <typename T>
int ft_fmexp(int w, int N, long long n, T* out, long long os, int omin,
             int omax, const T* mult, long long ms, int mmin, int mmax,
             const T* ex, long long es, int emin, int emax) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::ft_fmexp(
        Strat<T>{}, {}, batch<T>(out, os, bh.basis, omin, omax),
        batch<T>(mult, ms, bh.basis, mmin, mmax),
        batch<T>(ex, es, bh.basis, emin, emax), bh.basis,
        static_cast<Index>(n)));
}

template <typename T>
int ft_log(int w, int N, long long n, T* out, long long os, int omin, int omax,
           const T* arg, long long as, int amin, int amax) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::ft_log(
        Strat<T>{}, {}, batch<T>(out, os, bh.basis, omin, omax),
        batch<T>(arg, as, bh.basis, amin, amax), bh.basis,
        static_cast<Index>(n)));
}

template <typename T>
int antipode(int sign_by_degree, int w, int N, long long n, T* out, long long os, int omin, int omax, const T* arg,
             long long as, int amin, int amax) {
    BasisHolder bh(w, N);
    auto O = batch<T>(out, os, bh.basis, omin, omax);
    auto A = batch<T>(arg, as, bh.basis, amin, amax);
    if (sign_by_degree)
        return code_of(rpp::ops::tensor_antipode(Strat<T>{}, {}, O, A, bh.basis, static_cast<Index>(n)));
    return code_of(rpp::ops::tensor_reflect(Strat<T>{}, {}, O, A, bh.basis, static_cast<Index>(n)));
}

template <typename T>
int pairing(int w, int N, long long n, T* out, const T* f, long long fs,
            int fmin, int fmax, const T* x, long long xs, int xmin, int xmax) {
    BasisHolder bh(w, N);
    return code_of(rpp::ops::tensor_pairing(
        Strat<T>{}, {}, rpp::make_scalar_batch(out),
        batch<T>(f, fs, bh.basis, fmin, fmax),
        batch<T>(x, xs, bh.basis, xmin, xmax), bh.basis,
        static_cast<Index>(n)));
}

template <typename T, bool Left>
int adj(int w, int N, long long n, T* out, long long os, int omin, int omax,
        const T* op, long long ps, int pmin, int pmax, const T* arg,
        long long gs, int gmin, int gmax) {
    BasisHolder bh(w, N);
    auto O = batch<T>(out, os, bh.basis, omin, omax);
    auto P = batch<T>(op, ps, bh.basis, pmin, pmax);
    auto G = batch<T>(arg, gs, bh.basis, gmin, gmax);
    if constexpr (Left)
        return code_of(rpp::ops::ft_adj_lmul(Strat<T>{}, {}, O, P, G, bh.basis,
                                             static_cast<Index>(n)));
    else
        return code_of(rpp::ops::ft_adj_rmul(Strat<T>{}, {}, O, P, G, bh.basis,
                                             static_cast<Index>(n)));
}

// The linear operations over an arbitrary grading: a TensorBasis object
// built on the given degree offsets (the operations read nothing else of the
// basis, cpu/.../linalg/*.hpp), so Lie gradings go through the same code.
struct GradedHolder {
    std::vector<Index> degree_begin;
    TBasis basis;
    GradedHolder(const long long* start, int depth)
        : degree_begin(start, start + depth + 2),
          basis(static_cast<Degree>(0), static_cast<Degree>(depth), degree_begin.data()) {}
};

template <typename T>
int vector_assign(const long long* start, int N, long long n, T* out, long long os, int omin, int omax,
                  const T* arg, long long as, int amin, int amax) {
    GradedHolder g(start, N);
    return code_of(rpp::ops::vector_assign(Strat<T>{}, {}, batch<T>(out, os, g.basis, omin, omax),
                                           batch<T>(arg, as, g.basis, amin, amax), g.basis,
                                           static_cast<Index>(n)));
}

template <typename T>
int vector_add(const long long* start, int N, long long n, T* out, long long os, int omin, int omax,
               const T* lhs, long long ls, int lmin, int lmax, const T* rhs, long long rs, int rmin,
               int rmax, T alpha, T beta) {
    GradedHolder g(start, N);
    return code_of(rpp::ops::vector_add(Strat<T>{}, {}, batch<T>(out, os, g.basis, omin, omax),
                                        batch<T>(lhs, ls, g.basis, lmin, lmax),
                                        batch<T>(rhs, rs, g.basis, rmin, rmax), g.basis,
                                        static_cast<Index>(n), alpha, beta));
}

template <typename T>
int vector_inplace_add(const long long* start, int N, long long n, T* lhs, long long ls, int lmin, int lmax,
                       const T* rhs, long long rs, int rmin, int rmax, T alpha) {
    GradedHolder g(start, N);
    return code_of(rpp::ops::vector_inplace_add(Strat<T>{}, {}, batch<T>(lhs, ls, g.basis, lmin, lmax),
                                                batch<T>(rhs, rs, g.basis, rmin, rmax), g.basis,
                                                static_cast<Index>(n), alpha));
}

// op: 0 vector_scalar_multiply, 1 vector_set_constant, 2 tensor_set_identity,
// 3 tensor_add_identity.
template <typename T>
int vector_scalar(int op, const long long* start, int N, long long n, T* vec, long long vs, int vmin, int vmax,
                  T scalar) {
    GradedHolder g(start, N);
    auto V = batch<T>(vec, vs, g.basis, vmin, vmax);
    const auto count = static_cast<Index>(n);
    switch (op) {
        case 0: return code_of(rpp::ops::vector_scalar_multiply(Strat<T>{}, {}, V, g.basis, count, scalar));
        case 1: return code_of(rpp::ops::vector_set_constant(Strat<T>{}, {}, V, g.basis, count, scalar));
        case 2: return code_of(rpp::ops::tensor_set_identity(Strat<T>{}, {}, V, g.basis, count, scalar));
        case 3: return code_of(rpp::ops::tensor_add_identity(Strat<T>{}, {}, V, g.basis, count, scalar));
        default: return 3;
    }
}

// Shuffle family, through the reference's free functions.  kind as the
// oracle's orc_batch_st_fma: 0 st_mul, 1 st_fma, 2 st_inplace_fma (out == a).
template <typename T>
int st_fma(int kind, int w, int N, long long n, T* out, long long os, int omin, int omax, const T* a, long long as,
           int amin, int amax, const T* b, long long bs, int bmin, int bmax, const T* c, long long cs, int cmin,
           int cmax, T alpha, T beta) {
    BasisHolder bh(w, N);
    auto O = batch<T>(out, os, bh.basis, omin, omax);
    auto Bb = batch<T>(b, bs, bh.basis, bmin, bmax);
    auto Cb = batch<T>(c, cs, bh.basis, cmin, cmax);
    const auto count = static_cast<Index>(n);
    switch (kind) {
        case 0: return code_of(rpp::ops::st_mul(Strat<T>{}, {}, O, Bb, Cb, bh.basis, count, beta));
        case 1:
            return code_of(rpp::ops::st_fma(Strat<T>{}, {}, O, batch<T>(a, as, bh.basis, amin, amax), Bb, Cb, bh.basis,
                                            count, alpha, beta));
        case 2: return code_of(rpp::ops::st_inplace_fma(Strat<T>{}, {}, O, Bb, Cb, bh.basis, count, alpha, beta));
        default: return 3;
    }
}

// opcode as the oracle's orc_batch_shuffle2: 0 st_adj_mul(out, op=p, arg=q),
// 1 left_hs_mul(out, integrator=p, integrand=q, beta), 2 left_hs_adj_lmul,
// 3 left_hs_adj_rmul (out, integrator=p, integrand=q).  The two half-shuffle
// adjoints are applied through their operation objects, one object at a time:
// their free functions forward a beta that the cpu::single_thread operators
// do not take.
template <typename T>
int shuffle2(int opcode, int w, int N, long long n, T* out, long long os, int omin, int omax, const T* p,
             long long ps, int pmin, int pmax, const T* q, long long qs, int qmin, int qmax, T beta) {
    BasisHolder bh(w, N);
    const auto count = static_cast<Index>(n);
    auto O = batch<T>(out, os, bh.basis, omin, omax);
    auto P = batch<T>(p, ps, bh.basis, pmin, pmax);
    auto Q = batch<T>(q, qs, bh.basis, qmin, qmax);
    switch (opcode) {
        case 0: return code_of(rpp::ops::st_adj_mul(Strat<T>{}, {}, O, P, Q, bh.basis, count));
        case 1: return code_of(rpp::ops::left_hs_mul(Strat<T>{}, {}, O, P, Q, bh.basis, count, beta));
        case 2:
        case 3: {
            typename Strat<T>::Context ctx{};
            auto pack = rpp::basis::make_basis_pack(bh.basis);
            for (long long i = 0; i < n; ++i) {
                rpp::DenseTensorView<T*, TBasis> ov(out + i * os, bh.basis, static_cast<Degree>(omin),
                                                    static_cast<Degree>(omax));
                rpp::DenseTensorView<const T*, TBasis> pv(p + i * ps, bh.basis, static_cast<Degree>(pmin),
                                                          static_cast<Degree>(pmax));
                rpp::DenseTensorView<const T*, TBasis> qv(q + i * qs, bh.basis, static_cast<Degree>(qmin),
                                                          static_cast<Degree>(qmax));
                if (opcode == 2)
                    rpp::ops::LeftHSAdjLMul<Strat<T>>{ctx, pack}(ctx, ov, pv, qv);
                else
                    rpp::ops::LeftHSAdjRMul<Strat<T>>{ctx, pack}(ctx, ov, pv, qv);
            }
            return 0;
        }
        default: return 3;
    }
}

// Chen recursion as a reference user writes it: one ft_fmexp per increment
// over the whole batch, ping-pong buffers, exponent view (1,1) over a
// width+1 buffer whose slot 0 is unused.  Paths share one length L;
// increments are [n, L, width] row-major.
template <typename T>
int signature(int w, int N, long long n, long long L, const T* incr, T* out,
              int threads) {
    BasisHolder bh(w, N);
    const long long dim = bh.degree_begin[N + 1];
    return split(n, threads, [&](long long f, long long c) {
        std::vector<T> a(static_cast<size_t>(c * dim), T(0));
        std::vector<T> b(static_cast<size_t>(c * dim), T(0));
        std::vector<T> e(static_cast<size_t>(c * (w + 1)), T(0));
        for (long long i = 0; i < c; ++i) a[static_cast<size_t>(i * dim)] = T(1);
        T* cur = a.data();
        T* nxt = b.data();
        for (long long t = 0; t < L; ++t) {
            for (long long i = 0; i < c; ++i)
                for (int l = 0; l < w; ++l)
                    e[static_cast<size_t>(i * (w + 1) + 1 + l)] =
                        incr[((f + i) * L + t) * w + l];
            auto err = rpp::ops::ft_fmexp(
                Strat<T>{}, {}, batch<T>(nxt, dim, bh.basis, 0, N),
                batch<T>(static_cast<const T*>(cur), dim, bh.basis, 0, N),
                batch<T>(static_cast<const T*>(e.data()), w + 1, bh.basis, 1, 1),
                bh.basis, static_cast<Index>(c));
            if (!err) return code_of(err);
            std::swap(cur, nxt);
        }
        std::memcpy(out + f * dim, cur, sizeof(T) * static_cast<size_t>(c * dim));
        return 0;
    });
}

// Hall basis + L2T/T2L matrices built by the reference, CSC as returned.
struct LieData {
    std::vector<Index> pairs;        // 2 * true_size
    std::vector<Index> degree_begin; // depth + 2
    std::vector<double> l2t_data, t2l_data;
    std::vector<Index> l2t_idx, l2t_off, t2l_idx, t2l_off;
    long long l2t_outer = 0, l2t_inner = 0, t2l_outer = 0, t2l_inner = 0;
};

} // namespace

extern "C" {

#define REF_FT(T, SFX)                                                         \
    int ref_ft_mul_##SFX(int w, int N, long long n, T* out, long long os,      \
                         int omin, int omax, const T* lhs, long long ls,       \
                         int lmin, int lmax, const T* rhs, long long rs,       \
                         int rmin, int rmax, T beta, int threads) {            \
        return ft_mul<T>(w, N, n, out, os, omin, omax, lhs, ls, lmin, lmax,    \
                         rhs, rs, rmin, rmax, beta, threads);                  \
    }                                                                          \
    int ref_ft_fma_##SFX(int w, int N, long long n, T* out, long long os,      \
                         int omin, int omax, const T* a, long long as,         \
                         int amin, int amax, const T* b, long long bs,         \
                         int bmin, int bmax, const T* c, long long cs,         \
                         int cmin, int cmax, T alpha, T beta) {                \
        return ft_fma<T>(w, N, n, out, os, omin, omax, a, as, amin, amax, b,   \
                         bs, bmin, bmax, c, cs, cmin, cmax, alpha, beta);      \
    }                                                                          \
    int ref_ft_inplace_fma_##SFX(int variant, int w, int N, long long n, T* a, \
                                 long long as, int amin, int amax, const T* b, \
                                 long long bs, int bmin, int bmax, const T* c, \
                                 long long cs, int cmin, int cmax, T alpha,    \
                                 T beta) {                                     \
        return ft_inplace_fma<T>(variant, w, N, n, a, as, amin, amax, b, bs,   \
                                 bmin, bmax, c, cs, cmin, cmax, alpha, beta);  \
    }                                                                          \
    int ref_ft_inplace_mul_##SFX(int w, int N, long long n, T* lhs,            \
                                 long long ls, int lmin, int lmax,             \
                                 const T* rhs, long long rs, int rmin,         \
                                 int rmax, T beta) {                           \
        return ft_inplace_mul<T>(w, N, n, lhs, ls, lmin, lmax, rhs, rs, rmin,  \
                                 rmax, beta);                                  \
    }                                                                          \
    int ref_ft_exp_##SFX(int w, int N, long long n, T* out, long long os,      \
                         int omin, int omax, const T* arg, long long as,       \
                         int amin, int amax) {                                 \
        return ft_exp<T>(w, N, n, out, os, omin, omax, arg, as, amin, amax);   \
    }                                                                          \
    int ref_ft_fmexp_##SFX(int w, int N, long long n, T* out, long long os,    \
                           int omin, int omax, const T* m, long long ms,       \
                           int mmin, int mmax, const T* e, long long es,       \
                           int emin, int emax) {                               \
        return ft_fmexp<T>(w, N, n, out, os, omin, omax, m, ms, mmin, mmax, e, \
                           es, emin, emax);                                    \
    }                                                                          \
    int ref_ft_log_##SFX(int w, int N, long long n, T* out, long long os,      \
                         int omin, int omax, const T* arg, long long as,       \
                         int amin, int amax) {                                 \
        return ft_log<T>(w, N, n, out, os, omin, omax, arg, as, amin, amax);   \
    }                                                                          \
    int ref_tensor_pairing_##SFX(int w, int N, long long n, T* out,            \
                                 const T* f, long long fs, int fmin, int fmax, \
                                 const T* x, long long xs, int xmin,           \
                                 int xmax) {                                   \
        return pairing<T>(w, N, n, out, f, fs, fmin, fmax, x, xs, xmin, xmax); \
    }                                                                          \
    int ref_ft_adj_lmul_##SFX(int w, int N, long long n, T* out, long long os, \
                              int omin, int omax, const T* op, long long ps,   \
                              int pmin, int pmax, const T* arg, long long gs,  \
                              int gmin, int gmax) {                            \
        return adj<T, true>(w, N, n, out, os, omin, omax, op, ps, pmin, pmax,  \
                            arg, gs, gmin, gmax);                              \
    }                                                                          \
    int ref_ft_adj_rmul_##SFX(int w, int N, long long n, T* out, long long os, \
                              int omin, int omax, const T* op, long long ps,   \
                              int pmin, int pmax, const T* arg, long long gs,  \
                              int gmin, int gmax) {                            \
        return adj<T, false>(w, N, n, out, os, omin, omax, op, ps, pmin, pmax, \
                             arg, gs, gmin, gmax);                             \
    }                                                                          \
    int ref_tensor_antipode_##SFX(int sign, int w, int N, long long n, T* out, \
                                  long long os, int omin, int omax,            \
                                  const T* arg, long long as, int amin,        \
                                  int amax) {                                  \
        return antipode<T>(sign, w, N, n, out, os, omin, omax, arg, as, amin,  \
                           amax);                                              \
    }                                                                          \
    int ref_signature_##SFX(int w, int N, long long n, long long L,            \
                            const T* incr, T* out, int threads) {              \
        return signature<T>(w, N, n, L, incr, out, threads);                   \
    }                                                                          \
    int ref_st_fma_##SFX(int kind, int w, int N, long long n, T* out,          \
                         long long os, int omin, int omax, const T* a,         \
                         long long as, int amin, int amax, const T* b,         \
                         long long bs, int bmin, int bmax, const T* c,         \
                         long long cs, int cmin, int cmax, T alpha, T beta) {  \
        return st_fma<T>(kind, w, N, n, out, os, omin, omax, a, as, amin,      \
                         amax, b, bs, bmin, bmax, c, cs, cmin, cmax, alpha,    \
                         beta);                                                \
    }                                                                          \
    int ref_shuffle2_##SFX(int opcode, int w, int N, long long n, T* out,      \
                           long long os, int omin, int omax, const T* p,       \
                           long long ps, int pmin, int pmax, const T* q,       \
                           long long qs, int qmin, int qmax, T beta) {         \
        return shuffle2<T>(opcode, w, N, n, out, os, omin, omax, p, ps, pmin,  \
                           pmax, q, qs, qmin, qmax, beta);                     \
    }                                                                          \
    int ref_vector_assign_##SFX(const long long* start, int N, long long n,    \
                                T* out, long long os, int omin, int omax,      \
                                const T* arg, long long as, int amin,          \
                                int amax) {                                    \
        return vector_assign<T>(start, N, n, out, os, omin, omax, arg, as,     \
                                amin, amax);                                   \
    }                                                                          \
    int ref_vector_add_##SFX(const long long* start, int N, long long n,       \
                             T* out, long long os, int omin, int omax,         \
                             const T* lhs, long long ls, int lmin, int lmax,   \
                             const T* rhs, long long rs, int rmin, int rmax,   \
                             T alpha, T beta) {                                \
        return vector_add<T>(start, N, n, out, os, omin, omax, lhs, ls, lmin,  \
                             lmax, rhs, rs, rmin, rmax, alpha, beta);          \
    }                                                                          \
    int ref_vector_inplace_add_##SFX(const long long* start, int N,            \
                                     long long n, T* lhs, long long ls,        \
                                     int lmin, int lmax, const T* rhs,         \
                                     long long rs, int rmin, int rmax,         \
                                     T alpha) {                                \
        return vector_inplace_add<T>(start, N, n, lhs, ls, lmin, lmax, rhs,    \
                                     rs, rmin, rmax, alpha);                   \
    }                                                                          \
    int ref_vector_scalar_##SFX(int op, const long long* start, int N,         \
                                long long n, T* vec, long long vs, int vmin,   \
                                int vmax, T scalar) {                          \
        return vector_scalar<T>(op, start, N, n, vec, vs, vmin, vmax, scalar); \
    }

REF_FT(double, f64)
REF_FT(float, f32)

// Build the Hall basis and both matrices with the reference builders.
void* ref_lie_build(int width, int depth) {
    auto* d = new LieData();
    rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width),
                                     static_cast<Degree>(depth));
    auto lb = hall.to_lie_basis();
    const Index ts = lb.true_size();
    d->pairs.resize(static_cast<size_t>(2 * ts));
    for (Index i = 0; i < ts; ++i) {
        auto [l, r] = lb[i];
        d->pairs[static_cast<size_t>(2 * i)] = l;
        d->pairs[static_cast<size_t>(2 * i + 1)] = r;
    }
    d->degree_begin.assign(lb.degree_begin, lb.degree_begin + depth + 2);
    auto l2t = rpp::sparse::make_l2t_matrix<double>(lb);
    d->l2t_data.assign(l2t.data(), l2t.data() + l2t.nnz());
    d->l2t_idx.assign(l2t.indices(), l2t.indices() + l2t.nnz());
    d->l2t_off.assign(l2t.offsets(), l2t.offsets() + l2t.outer_dim() + 1);
    d->l2t_outer = l2t.outer_dim();
    d->l2t_inner = l2t.inner_dim();
    rpp::LieMultiplicationCache<Arch> cache(lb);
    auto t2l = rpp::sparse::make_t2l_matrix<double>(lb, cache);
    d->t2l_data.assign(t2l.data(), t2l.data() + t2l.nnz());
    d->t2l_idx.assign(t2l.indices(), t2l.indices() + t2l.nnz());
    d->t2l_off.assign(t2l.offsets(), t2l.offsets() + t2l.outer_dim() + 1);
    d->t2l_outer = t2l.outer_dim();
    d->t2l_inner = t2l.inner_dim();
    return d;
}

// sizes[0..5] = true_size, l2t nnz, l2t outer, l2t inner, t2l nnz, t2l outer,
// sizes[6] = t2l inner
void ref_lie_sizes(void* h, long long* sizes) {
    auto* d = static_cast<LieData*>(h);
    sizes[0] = static_cast<long long>(d->pairs.size() / 2);
    sizes[1] = static_cast<long long>(d->l2t_data.size());
    sizes[2] = d->l2t_outer;
    sizes[3] = d->l2t_inner;
    sizes[4] = static_cast<long long>(d->t2l_data.size());
    sizes[5] = d->t2l_outer;
    sizes[6] = d->t2l_inner;
}

void ref_lie_copy(void* h, long long* pairs, long long* degree_begin,
                  double* l2t_data, long long* l2t_idx, long long* l2t_off,
                  double* t2l_data, long long* t2l_idx, long long* t2l_off) {
    auto* d = static_cast<LieData*>(h);
    for (size_t i = 0; i < d->pairs.size(); ++i) pairs[i] = d->pairs[i];
    for (size_t i = 0; i < d->degree_begin.size(); ++i) degree_begin[i] = d->degree_begin[i];
    for (size_t i = 0; i < d->l2t_data.size(); ++i) {
        l2t_data[i] = d->l2t_data[i];
        l2t_idx[i] = d->l2t_idx[i];
    }
    for (size_t i = 0; i < d->l2t_off.size(); ++i) l2t_off[i] = d->l2t_off[i];
    for (size_t i = 0; i < d->t2l_data.size(); ++i) {
        t2l_data[i] = d->t2l_data[i];
        t2l_idx[i] = d->t2l_idx[i];
    }
    for (size_t i = 0; i < d->t2l_off.size(); ++i) t2l_off[i] = d->t2l_off[i];
}

void ref_lie_free(void* h) { delete static_cast<LieData*>(h); }

// One bracket from the reference LieMultiplicationCache: writes up to `cap`
// (key, coefficient) pairs and returns the count.
int ref_lie_bracket(int width, int depth, long long left, long long right,
                    long long* keys, long long* coeffs, int cap) {
    rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width),
                                     static_cast<Degree>(depth));
    rpp::LieMultiplicationCache<Arch> cache(hall.to_lie_basis());
    auto e = cache.get_bracket(static_cast<Index>(left), static_cast<Index>(right));
    int n = 0;
    for (auto const& [k, c] : e) {
        if (n < cap) {
            keys[n] = k;
            coeffs[n] = c;
        }
        ++n;
    }
    return n;
}

// tensor_to_lie / lie_to_tensor through the reference ops with the matrices
// the reference builds (CSC as built, or CSR via swap_format), FP64.
int ref_lie_convert_f64(int to_lie, int csr, int width, int depth, long long n,
                        double* out, long long os, const double* in,
                        long long is) {
    BasisHolder bh(width, depth);
    rpp::basis::HallBasis<Arch> hall(static_cast<Degree>(width),
                                     static_cast<Degree>(depth));
    auto lb = hall.to_lie_basis();
    const auto nn = static_cast<Index>(n);
    const auto D = static_cast<Degree>(depth);
    if (to_lie) {
        rpp::LieMultiplicationCache<Arch> cache(lb);
        auto m = rpp::sparse::make_t2l_matrix<double>(lb, cache);
        auto lie_out = rpp::make_lie_batch(out, static_cast<Index>(os), lb, Degree{1}, D);
        auto t_in = rpp::make_tensor_batch(in, static_cast<Index>(is), bh.basis, Degree{0}, D);
        if (csr) {
            auto m2 = rpp::sparse::swap_format(m.view());
            return code_of(rpp::ops::tensor_to_lie(Strat<double>{}, {}, lie_out, t_in,
                                                   lb, bh.basis, nn, m2.view()));
        }
        return code_of(rpp::ops::tensor_to_lie(Strat<double>{}, {}, lie_out, t_in, lb,
                                               bh.basis, nn, m.view()));
    }
    auto m = rpp::sparse::make_l2t_matrix<double>(lb);
    auto t_out = rpp::make_tensor_batch(out, static_cast<Index>(os), bh.basis, Degree{0}, D);
    auto lie_in = rpp::make_lie_batch(in, static_cast<Index>(is), lb, Degree{1}, D);
    if (csr) {
        auto m2 = rpp::sparse::swap_format(m.view());
        return code_of(rpp::ops::lie_to_tensor(Strat<double>{}, {}, t_out, lie_in,
                                               bh.basis, lb, nn, m2.view()));
    }
    return code_of(rpp::ops::lie_to_tensor(Strat<double>{}, {}, t_out, lie_in, bh.basis,
                                           lb, nn, m.view()));
}

} // extern "C"
