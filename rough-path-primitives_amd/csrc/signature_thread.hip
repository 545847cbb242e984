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
#include <mutex>

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
    // Increments one step ahead, in two register sets used in turn (the loop
    // is unrolled by two, so no copies): the row, uniform across the
    // wavefront, and the prefix letters' entries, loaded from the row pointer
    // with 32-bit lane offsets.
    T xa[W], xqa[L > 0 ? L : 1], xb[W], xqb[L > 0 ? L : 1];
    auto load = [&](T (&x)[W], T (&xq)[L > 0 ? L : 1], int64_t t) {
        const T* row = incr + t * W;
#pragma unroll
        for (int b = 0; b < W; ++b) x[b] = row[b];
#pragma unroll
        for (int i = 0; i < L; ++i) xq[i] = row[ql[i]];
    };
    // checkpoints: a countdown instead of a division per increment
    int ck_left = ck_every;
    int64_t ck_index = ck_every > 0 ? ck_offs[path] : 0;
    auto checkpoint = [&]() {
        if (ck_every > 0 && --ck_left == 0) {
            ck_left = ck_every;
            if (live) thread_store<T, W, N, L>(B, ckpts + ck_index * dim, 1, q, chain, sub);
            ++ck_index;
        }
    };
    if (t0 < t1) load(xa, xqa, t0);
    for (int64_t t = t0; t < t1; t += 2) {
        if (t + 1 < t1) load(xb, xqb, t + 1);
        thread_step<T, W, N, L>(chain, sub, xa, xqa);
        checkpoint();
        if (t + 1 >= t1) break;
        if (t + 2 < t1) load(xa, xqa, t + 2);
        thread_step<T, W, N, L>(chain, sub, xb, xqb);
        checkpoint();
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

// ------------------------------------------------------------------ backward
//
// One work-group per (path, first letter a); thread r owns the prefix
// q = a W^(L-1) + r of L letters.  With S = S_{t-1} and G = G_t (the
// cotangent of S_t = S (x) exp(x)), a step needs
//     Gbar[u] = sum_v G[u v] exp(x)[v]   and   xbar = d <G, S (x) exp(x)> / dx.
// Reversing the Horner recursion of every target word w (level K): Abar_K =
// G[w], and for i = K..1, xbar[w_i] += Abar_i A_{i-1} / (K-i+1),
// Gbar[w_1..w_i] += Abar_i, Abar_{i-1} = Abar_i x[w_i] / (K-i+1).  A thread
// takes the targets of its subtree and the chain words it owns (q with the
// remaining letters 0) and walks them down to level 1 on its own data:
//   - S_{t-1} is rebuilt from S_t by a Chen step with -x (on levels below N
//     only: level N of S never enters the backward step),
//   - the Horner values A come from S_{t-1} and x,
//   - Gbar of its subtree words is complete in the thread (every extension
//     of a subtree word lies in the same subtree) and replaces G in place,
//   - Gbar of a chain word (levels 1..L under a) collects one sum from every
//     thread below it: each thread writes its L sums to LDS and the block
//     sums over groups of W^(L-i) consecutive threads in L - 1 stages of W
//     terms, a fixed order; the result is the item's low-level cotangent for
//     the next step,
//   - xbar is summed within the work-group in a fixed order and written per
//     (increment, item); a second kernel adds the W items of each increment.
// With the S_0 cotangent unused (S_0 is the constant 1 along a path) no value
// crosses work-groups.  Padding threads (r >= W^(L-1)) contribute nothing.
template <int W, int N, int L>
struct BwdGeo {
    static constexpr int R = tpow(W, L - 1);        // threads with a prefix
    static constexpr int NT = (R + 63) / 64 * 64;  // padded to whole wavefronts
    // offset of level i (1..L) in the low-level block under a
    static constexpr int low_offset(int i) {
        int o = 0;
        for (int j = 1; j < i; ++j) o += tpow(W, j - 1);
        return o;
    }
    static constexpr int low_total = low_offset(L + 1);
};

template <typename T, int W, int N, int L, int NT = BwdGeo<W, N, L>::NT>
__global__ void __launch_bounds__(NT)
k_sig_bwd_thread(DevBasis B, int64_t n_items, const int64_t* order, const T* incr, const int64_t* offs,
                 View<const T> sigs, View<const T> cot, T* partial, int ck_every, const T* ckpts,
                 const int64_t* ck_offs) {
    static_assert(L >= 1 && L <= N - 2, "the prefix leaves a subtree below level N");
    using GS = ThreadGeo<W, N - 1, L>;  // S below level N
    using GG = ThreadGeo<W, N, L>;      // G, and the Horner values
    using BG = BwdGeo<W, N, L>;
    constexpr int R = BG::R;
    constexpr int NWV = NT / 64;
    __shared__ T glow[BG::low_total];
    __shared__ T cbuf[2][L > 1 ? L - 1 : 1][NT];
    __shared__ T red[NWV][W];
    const int64_t item = blockIdx.x;
    if (item >= n_items) return;
    const int64_t slot = item / W;
    const int a = static_cast<int>(item - slot * W);
    const int64_t path = order ? order[slot] : slot;
    const int t = threadIdx.x;
    const bool live = t < R;
    const int q = a * R + (live ? t : 0);
    int ql[L];
#pragma unroll
    for (int i = 0; i < L; ++i) ql[i] = (q / tpow(W, L - 1 - i)) % W;
    // S_t (below level N) and G_t
    T chain[L + 1];
    T subS[GS::total];
    T subG[GG::total];
    const T* s0 = sigs.obj(path);
    const T* g0 = cot.obj(path);
#pragma unroll
    for (int i = 0; i <= L; ++i) chain[i] = s0[(B.start[i] + q / tpow(W, L - i)) * sigs.cs];
#pragma unroll
    for (int i = L + 1; i < N; ++i)
#pragma unroll
        for (int rr = 0; rr < GS::words(i); ++rr)
            subS[GS::offset(i) + rr] = s0[(B.start[i] + static_cast<int64_t>(q) * GS::words(i) + rr) * sigs.cs];
#pragma unroll
    for (int i = L + 1; i <= N; ++i)
#pragma unroll
        for (int rr = 0; rr < GG::words(i); ++rr)
            subG[GG::offset(i) + rr] = g0[(B.start[i] + static_cast<int64_t>(q) * GG::words(i) + rr) * cot.cs];
    for (int j = t; j < BG::low_total; j += NT) {
        int lvl = 1;
#pragma unroll
        for (int i = 2; i <= L; ++i)
            if (j >= BG::low_offset(i)) lvl = i;
        const int u = j - BG::low_offset(lvl);
        glow[j] = g0[(B.start[lvl] + static_cast<int64_t>(a) * tpow(W, lvl - 1) + u) * cot.cs];
    }
    __syncthreads();
    const int64_t t0 = offs[path], t1 = offs[path + 1];
    const int64_t dim = B.start[N + 1];
    // increments one step ahead: the row (uniform) and the prefix letters'
    T xn[W], xqn[L];
    if (t0 < t1) {
#pragma unroll
        for (int b = 0; b < W; ++b) xn[b] = incr[(t1 - 1) * W + b];
#pragma unroll
        for (int i = 0; i < L; ++i) xqn[i] = incr[(t1 - 1) * W + ql[i]];
    }
    for (int64_t tt = t1 - 1; tt >= t0; --tt) {
        T x[W], xq[L], xneg[W], xqneg[L];
#pragma unroll
        for (int b = 0; b < W; ++b) {
            x[b] = xn[b];
            xneg[b] = -x[b];
        }
#pragma unroll
        for (int i = 0; i < L; ++i) {
            xq[i] = xqn[i];
            xqneg[i] = -xq[i];
        }
        if (tt > t0) {
#pragma unroll
            for (int b = 0; b < W; ++b) xn[b] = incr[(tt - 1) * W + b];
#pragma unroll
            for (int i = 0; i < L; ++i) xqn[i] = incr[(tt - 1) * W + ql[i]];
        }
        // S_{t-1}: identity before the first increment, a checkpoint where
        // one was stored, else S_t (x) exp(-x)
        const int64_t tl = tt - t0;
        if (tl == 0) {
            chain[0] = T(1);
#pragma unroll
            for (int i = 1; i <= L; ++i) chain[i] = T(0);
#pragma unroll
            for (int rr = 0; rr < GS::total; ++rr) subS[rr] = T(0);
        } else if (ck_every > 0 && (tl % ck_every) == 0) {
            const T* c0 = ckpts + (ck_offs[path] + tl / ck_every - 1) * dim;
#pragma unroll
            for (int i = 0; i <= L; ++i) chain[i] = c0[B.start[i] + q / tpow(W, L - i)];
#pragma unroll
            for (int i = L + 1; i < N; ++i)
#pragma unroll
                for (int rr = 0; rr < GS::words(i); ++rr)
                    subS[GS::offset(i) + rr] = c0[B.start[i] + static_cast<int64_t>(q) * GS::words(i) + rr];
        } else {
            thread_step<T, W, N - 1, L>(chain, subS, xneg, xqneg);
        }
        // Horner values of the step S_{t-1} -> S_t: A^(K)_L along q for the
        // subtree targets K > L, and A^(K)_i of subtree levels i < N
        T ach[N + 1];
#pragma unroll
        for (int K = L + 1; K <= N; ++K) {
            T v = chain[0];
#pragma unroll
            for (int i = 1; i <= L; ++i) v = (v * trecip<T>(K - i + 1)) * xq[i - 1] + chain[i];
            ach[K] = v;
        }
        T av[GG::atotal];
#pragma unroll
        for (int i = L + 1; i < N; ++i) {
#pragma unroll
            for (int u = 0; u < GG::words(i - 1); ++u) {
#pragma unroll
                for (int K = i + 1; K <= N; ++K) {
                    const T ap = (i == L + 1) ? ach[K] : av[GG::aoffset(i - 1) + (K - i) * GG::words(i - 1) + u];
                    const T apc = ap * trecip<T>(K - i + 1);
#pragma unroll
                    for (int b = 0; b < W; ++b)
                        av[GG::aoffset(i) + (K - i - 1) * GG::words(i) + u * W + b] =
                            apc * x[b] + subS[GS::offset(i) + u * W + b];
                }
            }
        }
        // reverse over the subtree, levels descending: ab holds Abar^(K)_i of
        // the level below the one processed (targets K > i), abl Abar^(K)_L
        T ab[GG::atotal];
#pragma unroll
        for (int j = 0; j < GG::atotal; ++j) ab[j] = T(0);
        T abl[N + 1];
#pragma unroll
        for (int K = 0; K <= N; ++K) abl[K] = T(0);
        T xbar[W];
#pragma unroll
        for (int b = 0; b < W; ++b) xbar[b] = T(0);
#pragma unroll
        for (int i = N; i > L; --i) {
#pragma unroll
            for (int u = 0; u < GG::words(i - 1); ++u) {
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const int rr = u * W + b;
                    T g = subG[GG::offset(i) + rr];  // Abar^(i)_i = G
                    T sum = g;
#pragma unroll
                    for (int K = i; K <= N; ++K) {
                        const T abk = (K == i) ? g : ab[GG::aoffset(i) + (K - i - 1) * GG::words(i) + rr];
                        if (K > i) sum += abk;
                        const T prev = (i == L + 1) ? ach[K] : av[GG::aoffset(i - 1) + (K - i) * GG::words(i - 1) + u];
                        const T abc = abk * trecip<T>(K - i + 1);
                        xbar[b] += abc * prev;
                        if (i == L + 1) {
                            abl[K] += abc * x[b];
                        } else {
                            ab[GG::aoffset(i - 1) + (K - i) * GG::words(i - 1) + u] += abc * x[b];
                        }
                    }
                    subG[GG::offset(i) + rr] = sum;
                }
            }
        }
        // the chain: subtree targets from level L, owned chain targets from
        // their own level; c[i] collects Gbar of the chain word of level i
        T c[L + 1];
#pragma unroll
        for (int i = 0; i <= L; ++i) c[i] = T(0);
        T xqbar[L];
#pragma unroll
        for (int i = 0; i < L; ++i) xqbar[i] = T(0);
#pragma unroll
        for (int K = N; K >= 1; --K) {
            const int top = K < L ? K : L;
            T abk;
            if (K > L) {
                abk = abl[K];
            } else {
                if (!live || q % tpow(W, L - K) != 0) continue;
                abk = glow[BG::low_offset(K) + q / tpow(W, L - K) - a * tpow(W, K - 1)];
            }
            T ca[L + 1];
            ca[0] = chain[0];
#pragma unroll
            for (int i = 1; i < top; ++i) ca[i] = (ca[i - 1] * trecip<T>(K - i + 1)) * xq[i - 1] + chain[i];
#pragma unroll
            for (int i = top; i >= 1; --i) {
                c[i] += abk;
                const T abc = abk * trecip<T>(K - i + 1);
                xqbar[i - 1] += abc * ca[i - 1];
                abk = abc * xq[i - 1];
            }
        }
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int b = 0; b < W; ++b)
                if (ql[i] == b) xbar[b] += xqbar[i];
        if (!live) {
#pragma unroll
            for (int i = 0; i <= L; ++i) c[i] = T(0);
#pragma unroll
            for (int b = 0; b < W; ++b) xbar[b] = T(0);
        }
        // xbar: wavefront sums, then the work-group's in wavefront order
#pragma unroll
        for (int b = 0; b < W; ++b) {
            T v = xbar[b];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if ((t & 63) == 0) red[t >> 6][b] = v;
        }
        // chain sums of levels < L to LDS (stage 0 of the block sums)
#pragma unroll
        for (int i = 1; i < L; ++i) cbuf[0][i - 1][t] = c[i];
        __syncthreads();  // every read of the old low-level cotangent is done
        if (live) glow[BG::low_offset(L) + t] = c[L];
        if (t < W) {
            T acc = T(0);
            for (int w = 0; w < NWV; ++w) acc += red[w][t];
            partial[(tt * W + a) * W + t] = acc;
        }
        // stage s sums groups of W entries of stage s - 1; after it the sums
        // of level L - s (groups of W^s threads) are complete
#pragma unroll
        for (int st = 1; st < L; ++st) {
            const int groups = tpow(W, L - 1 - st);
            if (t < groups) {
#pragma unroll
                for (int i = 1; i <= L - st; ++i) {
                    T acc = T(0);
#pragma unroll
                    for (int b = 0; b < W; ++b) acc += cbuf[(st - 1) & 1][i - 1][t * W + b];
                    if (i == L - st) {
                        glow[BG::low_offset(i) + t] = acc;
                    } else {
                        cbuf[st & 1][i - 1][t] = acc;
                    }
                }
            }
            __syncthreads();
        }
        if (L == 1) __syncthreads();
    }
}

// Partial rows are indexed by absolute increment row; rows outside
// [offsets[0], offsets[num_paths]) belong to no path, hold no partial sums and
// are left untouched in grad.
template <typename T, int W>
__global__ void k_sum_items(const T* partial, T* grad, int64_t total, const int64_t* offsets, int64_t num_paths) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total * W) return;
    const int64_t t = i / W;
    if (t < offsets[0] || t >= offsets[num_paths]) return;
    const int b = static_cast<int>(i - t * W);
    T acc = T(0);
    for (int a = 0; a < W; ++a) acc += partial[(t * W + a) * W + b];
    grad[i] = acc;
}

template <typename T, int W, int N, int L, int NT>
int thread_bwd_launch(const DevBasis& B, const rpp_hip_signature_backward_desc& d, const int64_t* order,
                      int64_t total_incr, hipStream_t stream, void* partial) {
    static_assert(BwdGeo<W, N, L>::NT <= 1024, "one work-group per (path, first letter)");
    const int64_t items = d.num_paths * W;
    if (items >= (int64_t{1} << 31)) return fail(RPP_HIP_NOT_IMPLEMENTED, "too many paths for one launch");
    void* part = partial;
    if (part == nullptr) {
        if (int rc = workspace(kSlotTmp, thread_backward_partial_bytes<T>(W, total_incr), &part)) return rc;
    }
    k_sig_bwd_thread<T, W, N, L><<<static_cast<unsigned>(items), BwdGeo<W, N, L>::NT, 0, stream>>>(
        B, items, order, static_cast<const T*>(d.increments), d.offsets, make_cview<T>(d.signatures),
        make_cview<T>(d.cotangent), static_cast<T*>(part), d.checkpoint_every, static_cast<const T*>(d.checkpoints),
        d.ckpt_offsets);
    if (int rc = check_hip(hipGetLastError(), "k_sig_bwd_thread launch")) return rc;
    if (total_incr > 0) {
        const int64_t n = total_incr * W;
        k_sum_items<T, W><<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
            static_cast<const T*>(part), static_cast<T*>(d.grad_increments), total_incr, d.offsets, d.num_paths);
    }
    return check_hip(hipGetLastError(), "k_sum_items launch");
}

}  // namespace

using ThreadBwdFn = int (*)(const DevBasis&, const rpp_hip_signature_backward_desc&, const int64_t*, int64_t,
                            hipStream_t, void*);

// the backward launcher of a tile, instantiated only for tiles that have one
template <typename T, int W, int N, int L, int NT, bool BWD>
constexpr ThreadBwdFn thread_bwd_of() {
    if constexpr (BWD) {
        return &thread_bwd_launch<T, W, N, L, NT>;
    } else {
        return nullptr;
    }
}

#define RPP_THREAD_TILE(T, W, N, L, NT, BWD) \
    {W, N, L, NT, &thread_fwd_launch<T, W, N, L, NT>, thread_bwd_of<T, W, N, L, NT, BWD>()}

template <typename T>
const std::vector<ThreadTile<T>>& thread_tiles() {
    static const std::vector<ThreadTile<T>> tiles = {
        // (width, depth, prefix letters per thread, forward work-group size,
        // with a backward); the backward's work-group is one (path, first
        // letter), W^(L-1) threads padded to whole wavefronts
        RPP_THREAD_TILE(T, 5, 6, 4, 256, 1), RPP_THREAD_TILE(T, 5, 6, 3, 256, 0),
        RPP_THREAD_TILE(T, 6, 6, 4, 256, 1),
        RPP_THREAD_TILE(T, 8, 5, 3, 256, 1),
        RPP_THREAD_TILE(T, 3, 8, 5, 256, 0), RPP_THREAD_TILE(T, 3, 8, 6, 256, 1),
    };
    return tiles;
}

template const std::vector<ThreadTile<float>>& thread_tiles<float>();
template const std::vector<ThreadTile<double>>& thread_tiles<double>();

// The allocation-free backward's scratch.  In host code `scratch` names a
// host-side stand-in, not device memory: only the address hipGetSymbolAddress
// returns for the current device is ever passed to a kernel.  Every call on
// one device shares that device's copy, which keeps the last call's contents.
namespace {
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(void** ptr) {
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(RPP_HIP_NOT_IMPLEMENTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return rc;
        if (p == nullptr) return fail(RPP_HIP_INTERNAL, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return 0;
}

}  // namespace rpp_hip
