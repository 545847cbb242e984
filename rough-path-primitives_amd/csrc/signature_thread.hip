// Thread-owned subtrees: the batched Chen recursion with one thread per
// (path, prefix q of L letters) and no communication at all.
//
// The Horner form of a Chen step S <- S (x) exp(x) for a word w = q v of
// level k > L,
//     A_0 = S_0,  A_i = A_{i-1} * x[w_i] / (k - i + 1) + S_i[w_1..w_i],  new S_k[w] = A_k,
// reads only S along the prefixes of w.  For i <= L those are the chain words
// S_0, S_1[q_1], .., S_L[q]; the rest lie in the subtree of q.  A thread that
// keeps a private copy of its chain (L + 1 values) and its subtree (the
// W + W^2 + .. + W^(N-L) words below q) in registers can therefore advance its
// part of the state by itself: no LDS, no barrier, no exchange.  The chain
// words are shared by the W^(L-i) threads below them, and each of those
// updates its own copy (new S_k[q_1..q_k] is the Horner value of target k
// along q), so the chain costs L(L+1)/2 + (N-L)L multiply-adds per thread and
// step on top of the subtree's sum_{i>L} (N-i+1) W^(i-L): at width 5 depth 6
// with L = 4, 36% more than the minimum, against a level-synchronous step's
// N - M barriers and LDS exchanges.
//
// Every value is computed from the old state before any is overwritten
// (chain and subtree targets read S_i before target i replaces it), so the
// step is the reference's truncated product S (x) exp(x) up to rounding.
//
// The threads of one path are padded to whole wavefronts (the padding lanes
// compute on prefix 0 and store nothing), so the path, its increments and
// the loop bounds are uniform across a wavefront.  Paths are taken in the
// queue order (longest first).  Stores: a thread writes its subtree, and the
// thread whose remaining prefix letters are all 0 writes each chain word.
#include "common.hpp"
#include "signature.hpp"
#include "signature_thread.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int tpow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// 1/m; m is fixed after unrolling at every use, so this folds to a literal
// (correctly rounded, like a runtime division).
template <typename T>
__device__ __forceinline__ T trecip(int m) {
    return T(1) / T(m);
}

template <int W, int N, int L>
struct ThreadGeo {
    static constexpr int P = tpow(W, L);              // prefixes (threads) per path
    static constexpr int PP = (P + 63) / 64 * 64;     // padded to whole wavefronts
    static constexpr int words(int i) { return tpow(W, i - L); }  // subtree words of level i >= L
    static constexpr int offset(int i) {              // register offset of subtree level i > L
        int o = 0;
        for (int j = L + 1; j < i; ++j) o += words(j);
        return o;
    }
    static constexpr int total = offset(N + 1);
    // Horner values A^(k)_i of levels i in [L+1, N-1] for targets k > i
    static constexpr int aoffset(int i) {
        int o = 0;
        for (int j = L + 1; j < i; ++j) o += (N - j) * words(j);
        return o;
    }
    static constexpr int atotal = aoffset(N) > 0 ? aoffset(N) : 1;
};

// One Chen step on a thread's chain copy and subtree.  x: the W increments;
// xq[i]: the increment of letter i + 1 of the prefix.
template <typename T, int W, int N, int L, int TOT = ThreadGeo<W, N, L>::total, int LQ = (L > 0 ? L : 1)>
__device__ __forceinline__ void thread_step(T (&chain)[L + 1], T (&sub)[TOT], const T (&x)[W], const T (&xq)[LQ]) {
    using G = ThreadGeo<W, N, L>;
    // chain Horner values from the old chain: A^(k)_L for the subtree targets
    // k > L, and the new chain words (targets k <= L)
    T ach[N + 1];
    T nch[L + 1];
#pragma unroll
    for (int k = 1; k <= N; ++k) {
        T a = chain[0];
#pragma unroll
        for (int i = 1; i <= L; ++i)
            if (i <= k) a = (a * trecip<T>(k - i + 1)) * xq[i - 1] + chain[i];
        if (k <= L) {
            nch[k] = a;
        } else {
            ach[k] = a;
        }
    }
    // subtree levels, ascending: level i forms A^(k)_i for every k >= i from
    // A^(k)_{i-1} and the old S_i; k = i is the new S_i
    T av[G::atotal];
#pragma unroll
    for (int i = L + 1; i <= N; ++i) {
#pragma unroll
        for (int u = 0; u < G::words(i - 1); ++u) {
            T apc[N + 1];
#pragma unroll
            for (int k = i; k <= N; ++k) {
                const T ap = (i == L + 1) ? ach[k] : av[G::aoffset(i - 1) + (k - i) * G::words(i - 1) + u];
                apc[k] = ap * trecip<T>(k - i + 1);
            }
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const T sv = sub[G::offset(i) + u * W + b];
#pragma unroll
                for (int k = i + 1; k <= N; ++k) av[G::aoffset(i) + (k - i - 1) * G::words(i) + u * W + b] = apc[k] * x[b] + sv;
                sub[G::offset(i) + u * W + b] = apc[i] * x[b] + sv;
            }
        }
    }
#pragma unroll
    for (int k = 1; k <= L; ++k) chain[k] = nch[k];
}

// Store a thread's subtree and the chain words it owns.
template <typename T, int W, int N, int L, int TOT = ThreadGeo<W, N, L>::total>
__device__ __forceinline__ void thread_store(const DevBasis& B, T* o, int64_t cs, int q, const T (&chain)[L + 1],
                                             const T (&sub)[TOT]) {
    using G = ThreadGeo<W, N, L>;
#pragma unroll
    for (int i = L + 1; i <= N; ++i) {
        T* lvl = o + (B.start[i] + static_cast<int64_t>(q) * G::words(i)) * cs;
#pragma unroll
        for (int r = 0; r < G::words(i); ++r) lvl[r * cs] = sub[G::offset(i) + r];
    }
#pragma unroll
    for (int i = 0; i <= L; ++i)
        if (q % tpow(W, L - i) == 0) o[(B.start[i] + q / tpow(W, L - i)) * cs] = chain[i];
}

template <typename T, int W, int N, int L, int NT>
__global__ void __launch_bounds__(NT)
k_sig_fwd_thread(DevBasis B, int n_paths, const int64_t* order, const T* incr, const int64_t* offs, View<T> out,
                 int ck_every, T* ckpts, const int64_t* ck_offs) {
    using G = ThreadGeo<W, N, L>;
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * NT + threadIdx.x;
    // a wavefront lies inside one path's padded range (NT and PP are
    // multiples of 64), so its path index is uniform
    const int slot = __builtin_amdgcn_readfirstlane(static_cast<int>(gid / G::PP));
    if (slot >= n_paths) return;
    const int lane_q = static_cast<int>(gid - static_cast<int64_t>(slot) * G::PP);
    const bool live = lane_q < G::P;
    const int q = live ? lane_q : 0;
    const int64_t path = order ? order[slot] : slot;
    const int64_t t0 = offs[path], t1 = offs[path + 1];
    int ql[L > 0 ? L : 1];  // letters of q, first (most significant) first
#pragma unroll
    for (int i = 0; i < L; ++i) ql[i] = (q / tpow(W, L - 1 - i)) % W;
    T chain[L + 1];
    chain[0] = T(1);
#pragma unroll
    for (int i = 1; i <= L; ++i) chain[i] = T(0);
    T sub[G::total];
#pragma unroll
    for (int r = 0; r < G::total; ++r) sub[r] = T(0);
    const int64_t dim = B.start[N + 1];
    // increments one step ahead: the whole row (uniform) and the prefix letters'
    T xn[W], xqn[L > 0 ? L : 1];
    if (t0 < t1) {
#pragma unroll
        for (int b = 0; b < W; ++b) xn[b] = incr[t0 * W + b];
#pragma unroll
        for (int i = 0; i < L; ++i) xqn[i] = incr[t0 * W + ql[i]];
    }
    for (int64_t t = t0; t < t1; ++t) {
        T x[W], xq[L > 0 ? L : 1];
#pragma unroll
        for (int b = 0; b < W; ++b) x[b] = xn[b];
#pragma unroll
        for (int i = 0; i < L; ++i) xq[i] = xqn[i];
        if (t + 1 < t1) {
#pragma unroll
            for (int b = 0; b < W; ++b) xn[b] = incr[(t + 1) * W + b];
#pragma unroll
            for (int i = 0; i < L; ++i) xqn[i] = incr[(t + 1) * W + ql[i]];
        }
        thread_step<T, W, N, L>(chain, sub, x, xq);
        if (ck_every > 0 && ((t - t0 + 1) % ck_every) == 0 && live)
            thread_store<T, W, N, L>(B, ckpts + (ck_offs[path] + (t - t0 + 1) / ck_every - 1) * dim, 1, q, chain, sub);
    }
    if (live) thread_store<T, W, N, L>(B, out.obj(path), out.cs, q, chain, sub);
}

template <typename T, int W, int N, int L, int NT>
int thread_fwd_launch(const DevBasis& B, const rpp_hip_signature_desc& d, const int64_t* order, hipStream_t stream) {
    using G = ThreadGeo<W, N, L>;
    static_assert(NT % 64 == 0 && G::PP % 64 == 0, "wavefront-aligned paths");
    if (d.num_paths >= (int64_t{1} << 31) / G::PP)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "too many paths for one thread-subtree launch");
    const int64_t threads = d.num_paths * G::PP;
    const unsigned grid = static_cast<unsigned>((threads + NT - 1) / NT);
    k_sig_fwd_thread<T, W, N, L, NT><<<grid, NT, 0, stream>>>(
        B, static_cast<int>(d.num_paths), order, static_cast<const T*>(d.increments), d.offsets, make_view<T>(d.out),
        d.checkpoint_every, static_cast<T*>(d.checkpoints), d.ckpt_offsets);
    return check_hip(hipGetLastError(), "k_sig_fwd_thread launch");
}

}  // namespace

#define RPP_THREAD_TILE(T, W, N, L, NT) {W, N, L, NT, &thread_fwd_launch<T, W, N, L, NT>}

template <typename T>
const std::vector<ThreadTile<T>>& thread_tiles() {
    static const std::vector<ThreadTile<T>> tiles = {
        // (width, depth, prefix letters per thread, work-group size)
        RPP_THREAD_TILE(T, 5, 6, 4, 256), RPP_THREAD_TILE(T, 5, 6, 3, 256),
        RPP_THREAD_TILE(T, 6, 6, 4, 256),
        RPP_THREAD_TILE(T, 8, 5, 3, 256),
        RPP_THREAD_TILE(T, 3, 8, 5, 256), RPP_THREAD_TILE(T, 3, 8, 6, 256),
    };
    return tiles;
}

template const std::vector<ThreadTile<float>>& thread_tiles<float>();
template const std::vector<ThreadTile<double>>& thread_tiles<double>();

}  // namespace rpp_hip
