// Truncated tensor products of small tensors (at most 2048 coefficients) with
// the shape fixed at compile time.
//
//     out_k[w] = beta * sum_j lhs_{k-j}[w / width^j] * rhs_j[w % width^j]
//                (+ alpha * add_k[w] on add's window)
//
// over the window-restricted j range, as product_pass in ft_ops.hip computes
// it: the same terms in the reference's CPU order (cpu_term), every product
// and sum rounded on its own (contraction off), so the results agree with the
// reference bit for bit.  With width and depth known to the compiler every
// degree offset, word count, division by width^j and the term order are
// constants, so the inner loops unroll into LDS reads, multiplies and adds.
// A work-group takes one object (a grid-stride loop covers batches beyond the
// grid limit; objects are taken in one contiguous range per XCD, 2^17 pairs
// at width 4 depth 4 FP64 194 -> 189.5 us): it stages the object's lhs and rhs windows in LDS with
// coalesced loads, then writes every output word once with coalesced stores.  All operand reads of an object precede its first write,
// so out may alias lhs or rhs; the addend is read at the index being written.
#include "common.hpp"
#include "product_small.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rpp_hip {

namespace {

constexpr int spow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

template <int W, int N>
struct SmallGeo {
    static constexpr int words(int k) { return spow(W, k); }
    static constexpr int start(int k) {
        int s = 0;
        for (int i = 0; i < k; ++i) s += spow(W, i);
        return s;
    }
    static constexpr int dim = start(N + 1);
};

// first word of degree k (k <= N + 1), by a loop over the compile-time degrees
template <int W, int N>
__device__ __forceinline__ int degree_start(int k) {
    int s = 0;
#pragma unroll
    for (int i = 0; i <= N; ++i)
        if (i < k) s += SmallGeo<W, N>::words(i);
    return s;
}

// FULL: every window is [0, N] and there is no addend, so the degree and
// split ranges are compile-time constants and the inner loops carry no
// branches (the general form tests each split against the windows).
template <typename T, int W, int N, int NT, bool FULL>
__global__ void __launch_bounds__(NT)
k_product_small(View<T> out, View<const T> lhs, View<const T> rhs, View<const T> add, int with_add, T alpha,
                T beta, int64_t n) {
    using G = SmallGeo<W, N>;
    __shared__ T sl[G::dim];
    __shared__ T sr[G::dim];
    const int t = threadIdx.x;
    const bool lw = lhs.mn <= lhs.mx, rw = rhs.mn <= rhs.mx;
    const int lb = lw ? degree_start<W, N>(lhs.mn) : 0, le = lw ? degree_start<W, N>(lhs.mx + 1) : 0;
    const int rb = rw ? degree_start<W, N>(rhs.mn) : 0, re = rw ? degree_start<W, N>(rhs.mx + 1) : 0;
    // Work-groups are dispatched to the eight XCDs round robin and the grid
    // is a multiple of 8, so v mod 8 names the XCD of virtual object v: XCD
    // x takes the contiguous objects [x per, (x + 1) per), and the cache
    // lines that neighbouring objects share meet in one L2.  The map is a
    // bijection of objects, so results never depend on the dispatch order.
    const int64_t per = (n + 7) >> 3;
    for (int64_t v = blockIdx.x; v < per * 8; v += gridDim.x) {
        const int64_t b = (v & 7) * per + (v >> 3);
        if (b >= n) continue;
        const T* L = lhs.obj(b);
        const T* R = rhs.obj(b);
        // compile-time trip count: every load of the object is issued before
        // the first LDS store waits on one
#pragma unroll
        for (int s = 0; s < (G::dim + NT - 1) / NT; ++s) {
            const int i = t + s * NT;
            if (G::dim % NT == 0 || i < G::dim) {
                if (FULL || (lb <= i && i < le)) sl[i] = stream_load(&L[i]);
                if (FULL || (rb <= i && i < re)) sr[i] = stream_load(&R[i]);
            }
        }
        __syncthreads();
        T* O = out.obj(b);
        const T* A = with_add ? add.obj(b) : nullptr;
        // the tiled order skips the scalar terms of a zero lhs[0] / rhs[0]
        const bool lhs0_zero = (FULL || lhs.mn == 0) && sl[0] == T(0);
        const bool rhs0_zero = (FULL || rhs.mn == 0) && sr[0] == T(0);
        // output degrees descending; k, each term's j and every index below
        // are compile-time constants
        static_for<N + 1>([&](auto ic) {
            constexpr int k = N - decltype(ic)::value;
            if (!FULL && (k < out.mn || k > out.mx)) return;
            const int jlo = FULL ? 0 : max(max(rhs.mn, k - lhs.mx), 0);
            const int jhi = FULL ? k : min(min(rhs.mx, k - lhs.mn), k);
            const bool use_add = !FULL && with_add && add.mn <= k && k <= add.mx;
#pragma unroll
            for (int s = 0; s < (G::words(k) + NT - 1) / NT; ++s) {
                const int w = t + s * NT;
                if (G::words(k) % NT == 0 || w < G::words(k)) {
                    constexpr int kF = cpu_first_tiled_degree(W), kR = cpu_right_letters(W);
                    constexpr bool tiled = W > 1 && k >= kF;
                    T acc = T(0);
                    static_for<k + 1>([&](auto uc) {
                        constexpr int j = cpu_term_fr(kF, kR, W == 1, k, decltype(uc)::value);
                        constexpr int wj = G::words(j), sk = G::start(k - j), sj = G::start(j);
                        if (tiled && j == k && lhs0_zero) return;
                        if (tiled && j == 0 && rhs0_zero) return;
                        if (FULL || (j >= jlo && j <= jhi)) {
                            const int p = w / wj;
                            acc += sl[sk + p] * sr[sj + (w - p * wj)];
                        }
                    });
                    T v = beta * acc;
                    if (use_add) v = alpha * A[G::start(k) + w] + v;
                    stream_store(v, &O[G::start(k) + w]);
                }
            }
        });
        __syncthreads();
    }
}

template <typename T, int W, int N>
int small_launch(int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& lhs,
                 const rpp_hip_tensor_batch& rhs, const rpp_hip_tensor_batch* add, double alpha, double beta,
                 hipStream_t stream) {
    constexpr int dim = SmallGeo<W, N>::dim;
    constexpr int NT = dim <= 128 ? 64 : (dim <= 1024 ? 128 : 256);
    // one object per work-group: the hardware scheduler keeps every compute
    // unit full and no work-group waits on a second object's loads
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(8, std::min<int64_t>((n + 7) / 8 * 8, 0x7ffffff8)));
    const View<const T> a = add ? make_cview<T>(*add) : View<const T>{nullptr, 0, 1, 1, 0};
    auto full = [](const rpp_hip_tensor_batch& t) { return t.min_degree == 0 && t.max_degree == N; };
    if (!add && full(out) && full(lhs) && full(rhs))
        k_product_small<T, W, N, NT, true><<<grid, NT, 0, stream>>>(
            make_view<T>(out), make_cview<T>(lhs), make_cview<T>(rhs), a, 0, T(alpha), T(beta), n);
    else
        k_product_small<T, W, N, NT, false><<<grid, NT, 0, stream>>>(
            make_view<T>(out), make_cview<T>(lhs), make_cview<T>(rhs), a, add ? 1 : 0, T(alpha), T(beta), n);
    return check_hip(hipGetLastError(), "k_product_small launch");
}

template <typename T>
using SmallFn = int (*)(int64_t, const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,
                        const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch*, double, double, hipStream_t);

template <typename T>
SmallFn<T> small_kernel(int width, int depth) {
#define RPP_SMALL(W, N) \
    if (width == W && depth == N) return &small_launch<T, W, N>;
    RPP_SMALL(2, 1) RPP_SMALL(2, 2) RPP_SMALL(2, 3) RPP_SMALL(2, 4) RPP_SMALL(2, 5) RPP_SMALL(2, 6)
    RPP_SMALL(2, 7) RPP_SMALL(2, 8) RPP_SMALL(2, 9)
    RPP_SMALL(3, 1) RPP_SMALL(3, 2) RPP_SMALL(3, 3) RPP_SMALL(3, 4) RPP_SMALL(3, 5) RPP_SMALL(3, 6)
    RPP_SMALL(4, 1) RPP_SMALL(4, 2) RPP_SMALL(4, 3) RPP_SMALL(4, 4) RPP_SMALL(4, 5)
    RPP_SMALL(5, 1) RPP_SMALL(5, 2) RPP_SMALL(5, 3) RPP_SMALL(5, 4)
    RPP_SMALL(6, 1) RPP_SMALL(6, 2) RPP_SMALL(6, 3) RPP_SMALL(6, 4)
    RPP_SMALL(8, 1) RPP_SMALL(8, 2) RPP_SMALL(8, 3)
#undef RPP_SMALL
    return nullptr;
}

}  // namespace

template <typename T>
bool product_small_available(int width, int depth) {
    return small_kernel<T>(width, depth) != nullptr;
}

template <typename T>
int launch_product_small(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                         const rpp_hip_tensor_batch& lhs, const rpp_hip_tensor_batch& rhs,
                         const rpp_hip_tensor_batch* add, double alpha, double beta, hipStream_t stream) {
    const SmallFn<T> fn = small_kernel<T>(B.width, B.depth);
    if (!fn) return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled small-tensor product for this shape");
    return fn(n, out, lhs, rhs, add, alpha, beta, stream);
}

template bool product_small_available<float>(int, int);
template bool product_small_available<double>(int, int);
template int launch_product_small<float>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                         const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,
                                         const rpp_hip_tensor_batch*, double, double, hipStream_t);
template int launch_product_small<double>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                          const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,
                                          const rpp_hip_tensor_batch*, double, double, hipStream_t);

}  // namespace rpp_hip
