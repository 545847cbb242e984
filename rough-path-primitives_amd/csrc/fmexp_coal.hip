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

// FULL: the multiplier window holds every degree, so its tests fold away.
// XCD: items are taken in XCD ranges (below); else item v is work-group v's.
// NTS: the subtree stores are non-temporal (the output is not read again
// before the caches have turned over).
template <typename T, int W, int N, int NT, int M, bool FULL, bool XCD, bool NTS>
__global__ void __launch_bounds__(NT)
k_fmexp_coal(DevBasis B, View<T> out, View<const T> mult_, View<const T> ex, int64_t n) {
    using G = CoalGeo<W, N, NT, M>;
    struct Mult {
        View<const T> v;
        __device__ const T* obj(int64_t b) const { return v.obj(b); }
        __device__ bool has(int k) const { return FULL || v.has(k); }
    };
    const Mult mult{mult_};
    const int64_t mcs = mult_.cs;
    __shared__ T lds0[G::lds_elems];
    __shared__ T lds1[G::lds_elems];
    __shared__ T xs[W];
    const int t = threadIdx.x;
    // Work-groups are dispatched to the eight XCDs round robin, so with a
    // grid that is a multiple of 8, v mod 8 names the XCD of virtual item v.
    // XCD x takes the contiguous item range [x per, (x + 1) per): the items
    // of one object share partial cache lines of its lower subtree levels
    // and its chain words, and these then meet in one XCD's L2.  The map is a
    // bijection of items, so results never depend on the dispatch order.
    const int64_t items = n * G::P;
    const int64_t per = XCD ? (items + 7) >> 3 : items;
    for (int64_t v = blockIdx.x; v < (XCD ? per * 8 : items); v += gridDim.x) {
        const int64_t item = XCD ? (v & 7) * per + (v >> 3) : v;
        if (XCD && item >= items) continue;
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
            chain[i] = mult.has(i) ? src[(B.start[i] + pre) * mcs] : T(0);
        }
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
            const bool in = mult.has(i);
            const T* lvl = src + (B.start[i] + static_cast<int64_t>(q) * G::words(i)) * mcs;
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                const int r = t + s * NT;
                if (G::words(i) % NT == 0 || r < G::words(i)) reg[G::offset(i) + s] = in ? lvl[r * mcs] : T(0);
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
                if (G::words(i) % NT == 0 || r < G::words(i)) {
                    if constexpr (NTS) {
                        stream_store(reg[G::offset(i) + s], &lvl[r * out.cs]);
                    } else {
                        lvl[r * out.cs] = reg[G::offset(i) + s];
                    }
                }
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

// Sibling-group wavefronts (tile kind 3).  A wavefront owns 64 consecutive
// sibling groups (the W children of one parent word) of one level k of one
// object.  Lane g forms the Horner value of target k along its group's
// parent p, A^(k)_{k-1}[p], from the chain words S_0, S_1[p_1], ..,
// S_{k-1}[p] read straight from memory (neighbouring lanes share most of
// them, so they come from the caches); the wavefront then writes the 64 W
// children in coalesced order, lane l taking word l + 64 s and its parent's
// value by a lane shuffle.  Every load and store of the object's own words
// covers 64 consecutive coefficients and no work-group barrier is needed:
// the exponent row is staged in a per-wavefront LDS slot, which replaces
// W + k - 1 lane-varying global reads per wavefront with LDS reads.  The
// roundings are horner_ref's in the order k_fmexp_coal uses, so both tiles
// equal the reference's ft_fmexp bit for bit.  Measured on MI355X against the
// subtree tile, both built from this file and timed in alternating rounds
// (microseconds per call, wavefront / subtree): FP64 width 6 depth 6 over 2048
// objects 322 / 332, width 8 depth 5 over 2048 objects 211 / 221; slower at
// width 5 depth 6 (252 / 235), width 3 depth 8, width 4 depth 4 and in FP32
// (width 8 depth 5 110 / 111, width 6 depth 6 226 / 175), where a
// wavefront's 64 groups hold fewer bytes for its fixed chain and staging work.
template <int W, int N>
struct WaveGeo {
    static constexpr int groups(int k) { return cpow(W, k - 1); }  // sibling groups of level k
    static constexpr int chunks(int k) { return (groups(k) + 63) / 64; }
    static constexpr int woff(int k) {
        int o = 0;
        for (int j = 1; j < k; ++j) o += chunks(j);
        return o;
    }
    static constexpr int per_object = woff(N + 1);  // wavefronts per object
    static constexpr int start(int k) {
        int o = 0;
        for (int j = 0; j < k; ++j) o += cpow(W, j);
        return o;
    }
};

template <typename T, int W, int N, int K>
__device__ __forceinline__ void wave_level(T* __restrict__ dst, const T* __restrict__ src, const T* xs, int mn,
                                           int mx, int c, int lane) {
    using G = WaveGeo<W, N>;
    auto has = [&](int i) { return mn <= i && i <= mx; };
    constexpr int NG = G::groups(K);
    const int g0 = c * 64;
    const int nwords = (NG - g0 < 64 ? NG - g0 : 64) * W;
    const int base = G::start(K) + g0 * W;
    T own[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
        const int j = lane + 64 * s;
        own[s] = (j < nwords && has(K)) ? src[base + j] : T(0);
    }
    // the lanes past the last group repeat it, so every read stays in range
    const int p = g0 + lane < NG ? g0 + lane : NG - 1;
    T pre[K], xl[K], xo[W];
    pre[0] = has(0) ? src[0] : T(0);
#pragma unroll
    for (int i = 1; i < K; ++i) {
        const int pi = p / cpow(W, K - 1 - i);
        pre[i] = has(i) ? src[G::start(i) + pi] : T(0);
        xl[i] = xs[pi % W];
    }
    // every exponent letter is read from LDS before the dependent steps
#pragma unroll
    for (int s = 0; s < W; ++s) xo[s] = xs[(lane + 64 * s) % W];
    T a = K == N ? pre[0] : pre[0] + T(0);
#pragma unroll
    for (int i = 1; i < K; ++i) a = horner_ref(a, xl[i], recip<T>(K - i + 1), pre[i], has(i));
#pragma unroll
    for (int s = 0; s < W; ++s) {
        const int j = lane + 64 * s;
        const T ap = __shfl(a, j / W, 64);
        if (j < nwords) stream_store(horner_ref(ap, xo[s], T(1), own[s], has(K)), &dst[base + j]);
    }
    if (K == 1 && lane == 0) dst[0] = pre[0] + T(0);  // the reference's last step leaves S_0 + 0
}

// One (object, level, chunk) per wavefront; unit coefficient strides; the
// output stores are non-temporal.  Work-groups take objects in launch order: taking them in one contiguous
// range per XCD, as k_fmexp_coal's FP64 width 5 depth 6 tile does, measured
// slower here (FP64 width 6 depth 6 over 2048 objects 323 -> 335 us per
// call, width 8 depth 5 212 -> 217).
template <typename T, int W, int N, int NT>
__global__ void __launch_bounds__(NT)
k_fmexp_wave(T* __restrict__ out, int64_t obs, const T* __restrict__ mult, int64_t mbs, int mn, int mx,
             const T* __restrict__ ex, int64_t ebs, int64_t n) {
    using G = WaveGeo<W, N>;
    static_assert(NT % 64 == 0, "whole wavefronts");
    __shared__ T xsl[NT / 64][W];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t gw = static_cast<int64_t>(blockIdx.x) * (NT / 64) + wv;
    const int64_t b = gw / G::per_object;
    if (b >= n) return;
    const int wl = static_cast<int>(gw - b * G::per_object);
    T* xs = xsl[wv];
    if (lane < W) xs[lane] = ex[b * ebs + 1 + lane];
    __builtin_amdgcn_wave_barrier();
    static_for<N>([&](auto I) {
        constexpr int K = decltype(I)::value + 1;
        if (wl >= G::woff(K) && wl < G::woff(K + 1))
            wave_level<T, W, N, K>(out + b * obs, mult + b * mbs, xs, mn, mx, wl - G::woff(K), lane);
    });
}

template <typename T, int W, int N, int NT>
int wave_launch(const DevBasis&, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& mult,
                const rpp_hip_tensor_batch& ex, int64_t n, hipStream_t stream) {
    using G = WaveGeo<W, N>;
    const int64_t blocks = (n * G::per_object + NT / 64 - 1) / (NT / 64);
    if (blocks >= (int64_t{1} << 31))
        return fail(RPP_HIP_NOT_IMPLEMENTED, "too many objects for one k_fmexp_wave launch");
    k_fmexp_wave<T, W, N, NT><<<static_cast<unsigned>(blocks), NT, 0, stream>>>(
        static_cast<T*>(out.data), out.batch_stride, static_cast<const T*>(mult.data), mult.batch_stride,
        mult.min_degree, mult.max_degree, static_cast<const T*>(ex.data), ex.batch_stride, n);
    return check_hip(hipGetLastError(), "k_fmexp_wave launch");
}

// FW: the tile also has a full-window instance, taken when the multiplier
// holds every degree (faster only at some shapes; see the tile table).
template <typename T, int W, int N, int NT, int M, int BPC, bool FW, bool XCD, bool NTS>
int coal_launch(const DevBasis& B, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& mult,
                const rpp_hip_tensor_batch& ex, int64_t n, hipStream_t stream) {
    using G = CoalGeo<W, N, NT, M>;
    // XCD: virtual items, the items rounded up to a multiple of 8 (one range
    // per XCD); a persistent grid is a multiple of 8 too (compute units x BPC)
    const int64_t virt = XCD ? (n * G::P + 7) / 8 * 8 : n * G::P;
    const int64_t cap = BPC > 0 ? static_cast<int64_t>(device_cu_count()) * BPC : (int64_t{1} << 30);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min(virt, cap)));
    if (FW && mult.min_degree <= 0 && mult.max_degree >= N)
        k_fmexp_coal<T, W, N, NT, M, true, XCD, NTS><<<grid, NT, 0, stream>>>(B, make_view<T>(out), make_cview<T>(mult),
                                                                    make_cview<T>(ex), n);
    else
        k_fmexp_coal<T, W, N, NT, M, false, XCD, NTS><<<grid, NT, 0, stream>>>(B, make_view<T>(out), make_cview<T>(mult),
                                                                     make_cview<T>(ex), n);
    return check_hip(hipGetLastError(), "k_fmexp_coal launch");
}

}  // namespace

#define RPP_COAL(T, W, N, NT, M, BPC, FW, XCD, NTS) \
    FmexpTile<T> { 2, sizeof(T), W, N, NT, M, BPC, &coal_launch<T, W, N, NT, M, BPC, FW, XCD, NTS> }
#define RPP_WAVE(T, W, N, NT) \
    FmexpTile<T> { 3, sizeof(T), W, N, NT, 0, 0, &wave_launch<T, W, N, NT> }

// One tile per (width, depth) serves plain ft_fmexp, the fastest of the
// candidates timed on MI355X.  Subtree tiles over 32 chained calls (threads,
// split: FP64 / FP32 microseconds per call):
//   width 5 depth 6, 4096 objects:  (128,2) 249 / 190, (64,2) 250 / 166, (256,1) 275 / 126
//   width 6 depth 6, 2048 objects:  (128,2) 338 / 179, (256,2) 341 / 181, (256,1) 414 / 283
//   width 8 depth 5, 2048 objects:  (256,1) 236 / 118, (256,2) 236 / 187, (128,2) 242 / 124
//   width 3 depth 8, 8192 objects:  (256,1) 262 / 146, (128,1) 268 / 174, (256,2) 315 / 279
//   width 4 depth 4, 2^17 objects:  (64,0) 143 / 75, (64,1) 213 / 189
// An FP32 item moves half the bytes of an FP64 one in the same number of
// dependent steps, so FP32 prefers fewer, larger items (one prefix letter).
// Sibling-group wavefront tiles (512 threads) replace them where they measured
// faster (k_fmexp_wave above): FP64 at width 6 depth 6 and width 8 depth 5.
// The full-window instance (last field), against the windowed one on a
// full-window multiplier, microseconds per call:
//   width 3 depth 8, 8192 objects:  FP64 252 -> 237, FP32 154 -> 134
//   width 6 depth 6, 2048 objects:  FP32 176 -> 168
//   width 5 depth 6, 4096 objects:  FP32 119 -> 114, FP64 237 -> 269 (not taken)
//   width 8 depth 5, 2048 objects:  FP32 111 -> 116 (not taken)
// Items in XCD ranges (XCD field), microseconds per call, alternating runs:
//   width 5 depth 6, 4096 objects:  FP64 235 -> 227 (taken); FP32 116 -> 118,
//   and no gain at the other shapes (not taken)
// Non-temporal subtree stores (last field; the wavefront tiles always), one
// box, alternating runs, microseconds per call:
//   FP64: width 4 depth 4 159 -> 134, width 3 depth 8 243 -> 238, width 6
//   depth 6 324 -> 319, width 8 depth 5 212 -> 207, width 5 depth 6
//   232 -> 238 (not taken; non-temporal for its top level only, 28.9 ->
//   30.7 ms per 128-call chain)
//   FP32: width 6 depth 6 199 -> 170, width 4 depth 4 82 -> 76, width 5
//   depth 6 126 -> 120, width 8 depth 5 117 -> 111, width 3 depth 8 147 -> 140
template <typename T>
const std::vector<FmexpTile<T>>& fmexp_coal_tiles() {
    constexpr bool f32 = sizeof(T) == 4;
    static const std::vector<FmexpTile<T>> tiles = {
        f32 ? RPP_COAL(T, 5, 6, 256, 1, 0, true, false, true) : RPP_COAL(T, 5, 6, 128, 2, 0, false, true, false),
        f32 ? RPP_COAL(T, 6, 6, 128, 2, 0, true, false, true) : RPP_WAVE(T, 6, 6, 512),
        f32 ? RPP_COAL(T, 8, 5, 256, 1, 0, false, false, true) : RPP_WAVE(T, 8, 5, 512),
        RPP_COAL(T, 3, 8, 256, 1, 0, true, false, true),
        RPP_COAL(T, 4, 4, 64, 0, 0, false, false, true),
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
