// Plain ft_fmexp with a degree-one exponent (one Chen step per call,
// out = mult (x) exp(x), full output window) as an HBM-streaming kernel.
//
// The call reads each multiplier once and writes each output once, so it is
// bound by HBM bandwidth.  A work item is (object b, prefix q of M letters):
// the subtree of every word of degree > M that starts with q, plus the chain
// words S_0 .. S_M along q (see signature_reg.hip for why subtrees advance
// independently).  Ownership is coalesced: thread t holds words
// r = t + s*NT of every subtree level in registers, so each wavefront load
// and store covers 64 consecutive coefficients.  The Horner scheme of target
// level k,
//     A_M = chain value,  A_i[r] = S_i[r] + (1 / (k - i + 1)) (0 + A_{i-1}[r / W] x[r % W]),
// is the reference's own sequence of roundings (horner_ref), so a call agrees
// with the reference's ft_fmexp bit for bit; it exchanges A_{i-1} through LDS
// (consecutive lanes read consecutive or equal parents: no bank conflicts),
// and the last level updates S_k in registers.  The chain starts from S_0 as
// the reference's step k + 1 leaves it: S_0 + 0 for k < N (its degree-0 tile
// is empty), S_0 itself for k = N.
#include "common.hpp"
#include "fmexp_coal.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int cpow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// 1/m; m is fixed after unrolling at every use, so this folds to a literal
// (correctly rounded, like a runtime division).
template <typename T>
__device__ __forceinline__ T recip(int m) {
    return T(1) / T(m);
}

template <int W, int N, int NT, int M>
struct CoalGeo {
    static constexpr int words(int i) { return cpow(W, i - M); }  // subtree words of level i > M
    static constexpr int slots(int i) { return (words(i) + NT - 1) / NT; }
    static constexpr int offset(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j) o += slots(j);
        return o;
    }
    static constexpr int total = offset(N + 1);
    // largest exchange: A_{N-1} over the subtree
    static constexpr int lds_elems = N - 1 > M ? words(N - 1) : 1;
    static constexpr int P = cpow(W, M);
};

template <typename T, int W, int N, int NT, int M>
__global__ void __launch_bounds__(NT)
k_fmexp_coal(DevBasis B, View<T> out, View<const T> mult, View<const T> ex, int64_t n) {
    using G = CoalGeo<W, N, NT, M>;
    __shared__ T lds0[G::lds_elems];
    __shared__ T lds1[G::lds_elems];
    __shared__ T xs[W];
    const int t = threadIdx.x;
    for (int64_t item = blockIdx.x; item < n * G::P; item += gridDim.x) {
        const int64_t b = item / G::P;
        const int q = static_cast<int>(item - b * G::P);
        const T* src = mult.obj(b);
        // the exponent row is read with the multiplier, so the two latencies
        // overlap instead of adding up across the barrier below
        const T xv = t < W ? ex.obj(b)[(1 + t) * ex.cs] : T(0);
        T reg[G::total > 0 ? G::total : 1];
        T chain[M + 1];
#pragma unroll
        for (int i = 0; i <= M; ++i) {
            const int pre = q / cpow(W, M - i);
            chain[i] = mult.has(i) ? src[(B.start[i] + pre) * mult.cs] : T(0);
        }
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
            const bool in = mult.has(i);
            const T* lvl = src + (B.start[i] + static_cast<int64_t>(q) * G::words(i)) * mult.cs;
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                const int r = t + s * NT;
                if (G::words(i) % NT == 0 || r < G::words(i)) reg[G::offset(i) + s] = in ? lvl[r * mult.cs] : T(0);
            }
        }
        __syncthreads();  // the previous item's readers of xs and LDS are done
        if (t < W) xs[t] = xv;
        __syncthreads();
#pragma unroll
        for (int k = N; k >= M + 1; --k) {
            T a = k == N ? chain[0] : chain[0] + T(0);
#pragma unroll
            for (int i = 1; i <= M; ++i)
                a = horner_ref(a, xs[(q / cpow(W, M - i)) % W], recip<T>(k - i + 1), chain[i], mult.has(i));
#pragma unroll
            for (int i = M + 1; i <= k; ++i) {
                const T c = recip<T>(k - i + 1);
                const T* prev = ((i - M) & 1) ? lds1 : lds0;
                T* cur = ((i - M) & 1) ? lds0 : lds1;
#pragma unroll
                for (int s = 0; s < G::slots(i); ++s) {
                    const int r = t + s * NT;
                    if (G::words(i) % NT == 0 || r < G::words(i)) {
                        const int p = r / W;
                        const int letter = r - p * W;
                        const T ap = (i == M + 1) ? a : prev[p];
                        const T v = horner_ref(ap, xs[letter], c, reg[G::offset(i) + s], mult.has(i));
                        if (i == k) {
                            reg[G::offset(i) + s] = v;
                        } else {
                            cur[r] = v;
                        }
                    }
                }
                if (i < k) __syncthreads();
            }
            __syncthreads();
        }
        // chain targets M .. 1, descending so each Horner reads old values
#pragma unroll
        for (int k = M; k >= 1; --k) {
            T a = chain[0] + T(0);
#pragma unroll
            for (int i = 1; i <= k; ++i)
                a = horner_ref(a, xs[(q / cpow(W, M - i)) % W], recip<T>(k - i + 1), chain[i], mult.has(i));
            chain[k] = a;
        }
        chain[0] = chain[0] + T(0);  // the reference's last step leaves S_0 + 0
        T* dst = out.obj(b);
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
            T* lvl = dst + (B.start[i] + static_cast<int64_t>(q) * G::words(i)) * out.cs;
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                const int r = t + s * NT;
                if (G::words(i) % NT == 0 || r < G::words(i)) lvl[r * out.cs] = reg[G::offset(i) + s];
            }
        }
        // chain word of level i is written by the item whose remaining
        // letters are all zero
        if (t == 0) {
#pragma unroll
            for (int i = 0; i <= M; ++i)
                if (q % cpow(W, M - i) == 0) dst[(B.start[i] + q / cpow(W, M - i)) * out.cs] = chain[i];
        }
    }
}

template <typename T, int W, int N, int NT, int M, int BPC>
int coal_launch(const DevBasis& B, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& mult,
                const rpp_hip_tensor_batch& ex, int64_t n, hipStream_t stream) {
    using G = CoalGeo<W, N, NT, M>;
    const int64_t items = n * G::P;
    const int64_t cap = BPC > 0 ? static_cast<int64_t>(device_cu_count()) * BPC : (int64_t{1} << 30);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min(items, cap)));
    k_fmexp_coal<T, W, N, NT, M><<<grid, NT, 0, stream>>>(B, make_view<T>(out), make_cview<T>(mult),
                                                          make_cview<T>(ex), n);
    return check_hip(hipGetLastError(), "k_fmexp_coal launch");
}

}  // namespace

#define RPP_COAL(T, W, N, NT, M, BPC) \
    FmexpTile<T> { sizeof(T), W, N, NT, M, BPC, &coal_launch<T, W, N, NT, M, BPC> }

// One tile per (width, depth) serves plain ft_fmexp, the fastest of the
// candidates timed on MI355X over 32 chained calls (threads, split: FP64 / FP32
// microseconds per call):
//   width 5 depth 6, 4096 objects:  (128,2) 249 / 190, (64,2) 250 / 166, (256,1) 275 / 126
//   width 6 depth 6, 2048 objects:  (128,2) 338 / 179, (256,2) 341 / 181, (256,1) 414 / 283
//   width 8 depth 5, 2048 objects:  (256,1) 236 / 118, (256,2) 236 / 187, (128,2) 242 / 124
//   width 3 depth 8, 8192 objects:  (256,1) 262 / 146, (128,1) 268 / 174, (256,2) 315 / 279
//   width 4 depth 4, 2^17 objects:  (64,0) 143 / 75, (64,1) 213 / 189
// An FP32 item moves half the bytes of an FP64 one in the same number of
// dependent steps, so FP32 prefers fewer, larger items (one prefix letter).
template <typename T>
const std::vector<FmexpTile<T>>& fmexp_coal_tiles() {
    constexpr bool f32 = sizeof(T) == 4;
    static const std::vector<FmexpTile<T>> tiles = {
        f32 ? RPP_COAL(T, 5, 6, 256, 1, 0) : RPP_COAL(T, 5, 6, 128, 2, 0),
        RPP_COAL(T, 6, 6, 128, 2, 0),
        RPP_COAL(T, 8, 5, 256, 1, 0),
        RPP_COAL(T, 3, 8, 256, 1, 0),
        RPP_COAL(T, 4, 4, 64, 0, 0),
    };
    return tiles;
}

template <typename T>
const FmexpTile<T>* fmexp_coal_tile(int width, int depth) {
    for (const auto& t : fmexp_coal_tiles<T>())
        if (t.width == width && t.depth == depth) return &t;
    return nullptr;
}

template const std::vector<FmexpTile<float>>& fmexp_coal_tiles<float>();
template const std::vector<FmexpTile<double>>& fmexp_coal_tiles<double>();
template const FmexpTile<float>* fmexp_coal_tile<float>(int, int);
template const FmexpTile<double>* fmexp_coal_tile<double>(int, int);

}  // namespace rpp_hip
