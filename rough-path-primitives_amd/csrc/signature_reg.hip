// Register-resident batched Chen recursion: the same Horner scheme as
// signature.hip, with each path's state held in VGPRs across one work-group
// and the Horner intermediates exchanged through LDS.
//
// Tile (compile time): width W, depth N, NT threads, and M leading letters
// fixed per work-group.  A work item is (path, prefix q of M letters); it
// owns every word of degree > M that starts with q (the "subtree"), plus the
// single word q of degree M (the "root") and S_0.  A Chen step never mixes
// subtrees: A_i[q u] depends only on A_{i-1}[q u'] and the chain
// A_0 .. A_M along q, so the subtrees of one path advance independently.
//
// Ownership inside a subtree: the words of level i > M are grouped by their
// parent (a word of level i-1); parent p is owned by thread p mod NT, which
// holds all W children p*W + b in consecutive registers.  One LDS read of
// A_{i-1}[p] feeds W fused multiply-adds.
//
// Level-synchronous schedule: the Horner recursions of all target levels k
// advance together, one subtree level per phase.  Phase i forms A^(k)_i for
// every k >= i from A^(k)_{i-1} (LDS) and the old S_i; k = i is the new S_i
// (in place, after the k > i values have read the old one), k > i go to LDS
// for phase i+1.  A Chen step then costs N - M barriers instead of one per
// (target, level) pair.
//
// Backward (M <= 1): S and the cotangent G share the layout.  Per increment:
// rebuild the previous state (a Chen step with -x), recompute the Horner
// values A^(K)_i of every target K (kept in LDS), then one reverse phase per
// level, descending: Abar^(K)_i of every live target arrives through LDS,
// contributes to xbar and to G_i, and Abar^(K)_{i-1} = c sum_b Abar^(K)_i[pb] x_b
// goes back to LDS.  All targets read the cotangent of their own top level
// before any lower target adds to it, exactly as in the target-by-target
// order (target K only adds into levels below K).
#include "common.hpp"
#include "signature.hpp"
#include "signature_reg.hpp"

#include <algorithm>
#include <tuple>
#include <utility>

namespace rpp_hip {

namespace {

constexpr int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// 1/m.  Every use has m fixed after unrolling, so the division folds to a
// literal, correctly rounded like a runtime division: no table, no load.
template <typename T>
__device__ __forceinline__ T recip(int m) {
    return T(1) / T(m);
}

template <int W, int N, int NT, int M, int RL>
struct Geo {
    // parents of level i (> M) inside one subtree: words of level i-1 under q
    static constexpr int parents(int i) { return ipow(W, i - 1 - M); }
    static constexpr int slots(int i) { return (parents(i) + NT - 1) / NT; }
    // The top RL levels live in registers; levels M+1 .. N-RL live in LDS,
    // where each word is read and written only by its owning thread.
    static constexpr bool in_reg(int i) { return i > N - RL; }
    // register offset of level i's first slot (register levels)
    static constexpr int offset(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j)
            if (in_reg(j)) o += slots(j) * W;
        return o;
    }
    static constexpr int total = offset(N + 1);
    // LDS offset of level i (LDS levels), laid out [letter][parent] so that
    // consecutive threads touch consecutive words
    static constexpr int state_offset(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j)
            if (!in_reg(j)) o += ipow(W, j - M);
        return o;
    }
    static constexpr int state_elems = state_offset(N + 1) > 0 ? state_offset(N + 1) : 1;
    // Occupancy the register allocator aims for (waves per SIMD): tiles that
    // keep lower levels in LDS are small enough in registers to run several
    // work-groups per compute unit; 8-byte values take twice the registers.
    static constexpr int fwd_waves(int bytes) { return RL < N - M ? (bytes == 8 ? 3 : 4) : 1; }
    static constexpr int bwd_waves(int bytes) { return RL < N - M ? (bytes == 8 ? 2 : 3) : 1; }
    static constexpr int apar_offset(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j) o += slots(j);
        return o;
    }
    static constexpr int apar_total = apar_offset(N + 1);
    // words of level i inside one subtree
    static constexpr int words(int i) { return ipow(W, i - M); }
    // largest LDS exchange: A^(k)_i for the live targets k > i of one level,
    // (N - i) * words(i), which peaks at i = N - 1 (= parents(N))
    static constexpr int lds_elems = N > M ? parents(N) : 1;
    // backward: A^(K)_i for i in [M+1, N-1], K in (i, N]
    static constexpr int astore_offset(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j) o += (N - j) * words(j);
        return o;
    }
    static constexpr int astore_elems = N - 1 > M ? astore_offset(N) : 1;
    static constexpr int P = ipow(W, M);
    // global word index of subtree word (level i, local r) for prefix q
    __device__ __forceinline__ static int64_t global_index(const DevBasis& B, int i, int64_t q, int64_t r) {
        return B.start[i] + q * static_cast<int64_t>(ipow(W, i - M)) + r;
    }
};

template <typename T, int W, int N, int NT, int M, int RL>
struct Tile {
    using G = Geo<W, N, NT, M, RL>;
    T reg[G::total > 0 ? G::total : 1];
    T* lst;  // this work-group's LDS block of the levels below the register levels
    T root;  // S_M[q] (for M = 0 this is S_0)
    T s0;    // S_0
    T mid[M > 1 ? M - 1 : 1];  // S_1[q_1] .. S_{M-1}[q_1..q_{M-1}] when M >= 2

    __device__ explicit Tile(T* lds_state) : lst(lds_state) {}

    // child b of owned parent slot s at level i (compile-time i, s, b)
    __device__ __forceinline__ T& word(int i, int s, int b) {
        if (G::in_reg(i)) return reg[G::offset(i) + s * W + b];
        return lst[G::state_offset(i) + b * G::parents(i) + threadIdx.x + s * NT];
    }
    __device__ __forceinline__ T word(int i, int s, int b) const {
        if (G::in_reg(i)) return reg[G::offset(i) + s * W + b];
        return lst[G::state_offset(i) + b * G::parents(i) + threadIdx.x + s * NT];
    }

    // prefix of q of length i (a word of level i), i <= M
    __device__ __forceinline__ static int prefix(int q, int i) { return q / ipow(W, M - i); }
    __device__ __forceinline__ static int letter(int q, int i) { return prefix(q, i) % W; }
    // chain value of level i (0 <= i <= M)
    // (for M = 0, root and s0 both hold S_0)
    __device__ __forceinline__ T chain(int i) const { return i == M ? root : (i == 0 ? s0 : mid[i - 1]); }
    __device__ __forceinline__ void set_chain(int i, T v) {
        if (i == M) root = v;
        if (i == 0) s0 = v;
        if (i != 0 && i != M) mid[i - 1] = v;
    }

    __device__ __forceinline__ void set_identity() {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                if (t + s * NT < G::parents(i)) {
#pragma unroll
                    for (int b = 0; b < W; ++b) word(i, s, b) = T(0);
                }
            }
        }
        root = M == 0 ? T(1) : T(0);
        s0 = T(1);
#pragma unroll
        for (int i = 0; i < (M > 1 ? M - 1 : 1); ++i) mid[i] = T(0);
    }

    // Store (out = true) or load the subtree words and the chain words this
    // work item owns.  Object base `o`, coefficient stride cs.
    template <bool STORE>
    __device__ __forceinline__ void move(const DevBasis& B, T* o, int64_t cs, int q, bool owns_chain) {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                const int p = t + s * NT;
                if (p < G::parents(i)) {
#pragma unroll
                    for (int b = 0; b < W; ++b) {
                        const int64_t gi = G::global_index(B, i, q, static_cast<int64_t>(p) * W + b);
                        if constexpr (STORE) {
                            o[gi * cs] = word(i, s, b);
                        } else {
                            word(i, s, b) = o[gi * cs];
                        }
                    }
                }
            }
        }
        // chain words: level i <= M holds the prefix of q of length i; it is
        // stored by the item whose remaining letters are all zero
        if constexpr (STORE) {
            if (t == 0) {
#pragma unroll
                for (int i = 0; i <= M; ++i)
                    if (q % ipow(W, M - i) == 0) o[(B.start[i] + prefix(q, i)) * cs] = chain(i);
            }
        } else {
#pragma unroll
            for (int i = 0; i <= M; ++i) set_chain(i, o[(B.start[i] + prefix(q, i)) * cs]);
        }
    }

    // Load the words of degree [mn, mx] (zero elsewhere), object-major.
    __device__ __forceinline__ void load_window(const DevBasis& B, const T* o, int q, int mn, int mx) {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = M + 1; i <= N; ++i) {
            const bool in = mn <= i && i <= mx;
#pragma unroll
            for (int s = 0; s < G::slots(i); ++s) {
                const int p = t + s * NT;
                if (p < G::parents(i)) {
#pragma unroll
                    for (int b = 0; b < W; ++b)
                        word(i, s, b) =
                            in ? o[G::global_index(B, i, q, static_cast<int64_t>(p) * W + b)] : T(0);
                }
            }
        }
#pragma unroll
        for (int i = 0; i <= M; ++i) set_chain(i, (mn <= i && i <= mx) ? o[B.start[i] + prefix(q, i)] : T(0));
    }

};

// Chain values A^(k)_M along the prefix q for every target k > M (uniform):
// ach[k] = Horner of S_0 .. S_M with x[q_i] / (k - i + 1).
template <typename T, int W, int N, int NT, int M, int RL>
__device__ __forceinline__ void chain_values(const Tile<T, W, N, NT, M, RL>& S, const T* xs, int q,
                                             T (&ach)[N + 1]) {
#pragma unroll
    for (int k = M + 1; k <= N; ++k) {
        T a = S.s0;
#pragma unroll
        for (int i = 1; i <= M; ++i) a = (a * recip<T>(k - i + 1)) * xs[S.letter(q, i)] + S.chain(i);
        if (M == 0) a = S.root;
        ach[k] = a;
    }
}

// One in-place Chen step S <- S (x) exp(x) on the tile, level-synchronous
// (see the file comment).  xs: W increments (uniform, unchanged until the
// step returns), lds: two buffers of G::lds_elems; phase i
// reads A^(k)_{i-1} at prev[(k - i) * parents(i) + p] and writes A^(k)_i at
// cur[(k - i - 1) * words(i) + p*W + b].
template <typename T, int W, int N, int NT, int M, int RL>
__device__ __forceinline__ void chen_step_reg(Tile<T, W, N, NT, M, RL>& S, const T* xs, T* lds0,
                                              T* lds1, int q) {
    using G = Geo<W, N, NT, M, RL>;
    const int t = threadIdx.x;
    T ach[N + 1];
    chain_values<T, W, N, NT, M, RL>(S, xs, q, ach);
    T xr[W];
#pragma unroll
    for (int b = 0; b < W; ++b) xr[b] = xs[b];
#pragma unroll
    for (int i = M + 1; i <= N; ++i) {
        const T* prev = ((i - M) & 1) ? lds1 : lds0;
        T* cur = ((i - M) & 1) ? lds0 : lds1;
#pragma unroll
        for (int s = 0; s < G::slots(i); ++s) {
            const int p = t + s * NT;
            if (p < G::parents(i)) {
#pragma unroll
                for (int k = N; k >= i; --k) {
                    // the parent value scaled once, then W fused multiply-adds
                    const T ap = (i == M + 1) ? ach[k] : prev[(k - i) * G::parents(i) + p];
                    const T apc = ap * recip<T>(k - i + 1);
#pragma unroll
                    for (int b = 0; b < W; ++b) {
                        T& sv = S.word(i, s, b);
                        const T v = apc * xr[b] + sv;
                        if (k == i) {
                            sv = v;
                        } else {
                            cur[(k - i - 1) * G::words(i) + p * W + b] = v;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    // target levels M .. 1: the chain words, descending so that each Horner
    // reads old lower chain values
#pragma unroll
    for (int k = M; k >= 1; --k) {
        T a = S.s0;
#pragma unroll
        for (int i = 1; i <= k; ++i) a = (a * recip<T>(k - i + 1)) * xs[S.letter(q, i)] + S.chain(i);
        S.set_chain(k, a);
    }
}

template <typename T, int W, int N, int NT, int M, int RL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Geo<W, N, NT, M, RL>::fwd_waves(sizeof(T)))))
k_sig_fwd_reg(DevBasis B, int64_t n_items, const int64_t* order, const T* incr, const int64_t* offs,
              View<T> out, int ck_every, T* ckpts, const int64_t* ck_offs, unsigned* counter) {
    using G = Geo<W, N, NT, M, RL>;
    __shared__ T lds0[G::lds_elems];
    __shared__ T lds1[G::lds_elems];
    __shared__ T sstate[G::state_elems];
    // increments, double-buffered by parity: a step still reading increment t
    // (its chain update follows the last barrier) never sees t + 1 arrive
    __shared__ T xs[2][W];
    __shared__ unsigned slot;
    const int64_t dim = B.start[N + 1];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) slot = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t item = slot;
        if (item >= n_items) break;
        const int64_t path = order ? order[item / G::P] : item / G::P;
        const int q = static_cast<int>(item % G::P);
        Tile<T, W, N, NT, M, RL> S(sstate);
        S.set_identity();
        const int64_t t0 = offs[path], t1 = offs[path + 1];
        // the next increment is loaded one step ahead, so its global-memory
        // latency overlaps the current step instead of stalling the barrier
        T xnext = (threadIdx.x < W && t0 < t1) ? incr[t0 * W + threadIdx.x] : T(0);
        for (int64_t t = t0; t < t1; ++t) {
            T* xt = xs[(t - t0) & 1];
            if (threadIdx.x < W) {
                xt[threadIdx.x] = xnext;
                if (t + 1 < t1) xnext = incr[(t + 1) * W + threadIdx.x];
            }
            __syncthreads();
            chen_step_reg<T, W, N, NT, M, RL>(S, xt, lds0, lds1, q);
            if (ck_every > 0 && ((t - t0 + 1) % ck_every) == 0) {
                T* dst = ckpts + (ck_offs[path] + (t - t0 + 1) / ck_every - 1) * dim;
                S.template move<true>(B, dst, 1, q, q == 0);
            }
        }
        S.template move<true>(B, out.obj(path), out.cs, q, q == 0);
    }
}

// ------------------------------------------------------------------ backward

// Reverse mode of one Chen step taken at state S (the state before the step),
// level-synchronous (see the file comment).  Gt, the cotangent of the new
// state, becomes in place the cotangent of S; xbar[b] accumulates this
// thread's share of the x gradient; thread 0 accumulates the chain terms: xq
// (to x[q], M = 1) and s0bar (the cotangent of S_0).  The cotangent of the
// root word (M = 1) and of target level M itself are the caller's: it seeds xq
// and s0bar from the old G_M before this call.
//   astore: A^(K)_i, i in [M+1, N-1], K in (i, N], at
//           astore_offset(i) + (K - i - 1) * words(i) + p*W + b;
//   lds0/lds1: Abar^(K)_{i-1}, K in [i, N], at (K - i) * parents(i) + p.
template <typename T, int W, int N, int NT, int M, int RL>
__device__ __forceinline__ void bwd_step(const Tile<T, W, N, NT, M, RL>& S, Tile<T, W, N, NT, M, RL>& Gt, const T* xs,
                                         T* astore, T* lds0, T* lds1, int q, T (&xbar)[W], T& xq,
                                         T& s0bar) {
    using G = Geo<W, N, NT, M, RL>;
    const int t = threadIdx.x;
    T ach[N + 1];
    chain_values<T, W, N, NT, M, RL>(S, xs, q, ach);
    T xr[W];
#pragma unroll
    for (int b = 0; b < W; ++b) xr[b] = xs[b];
    // forward Horner values of every target, levels M+1 .. N-1
#pragma unroll
    for (int i = M + 1; i <= N - 1; ++i) {
#pragma unroll
        for (int s = 0; s < G::slots(i); ++s) {
            const int p = t + s * NT;
            if (p < G::parents(i)) {
#pragma unroll
                for (int K = N; K > i; --K) {
                    const T ap =
                        (i == M + 1) ? ach[K] : astore[G::astore_offset(i - 1) + (K - i) * G::parents(i) + p];
                    const T apc = ap * recip<T>(K - i + 1);
#pragma unroll
                    for (int b = 0; b < W; ++b)
                        astore[G::astore_offset(i) + (K - i - 1) * G::words(i) + p * W + b] =
                            apc * xr[b] + S.word(i, s, b);
                }
            }
        }
        __syncthreads();
    }
    // reverse phases, levels N .. M+1
#pragma unroll
    for (int i = N; i >= M + 1; --i) {
        const T* src = ((N - i) & 1) ? lds1 : lds0;  // Abar^(K)_i, K in (i, N]
        T* dst = ((N - i) & 1) ? lds0 : lds1;        // Abar^(K)_{i-1}, K in [i, N]
#pragma unroll
        for (int s = 0; s < G::slots(i); ++s) {
            const int p = t + s * NT;
            if (p < G::parents(i)) {
                T gsum[W];
#pragma unroll
                for (int b = 0; b < W; ++b) gsum[b] = T(0);
#pragma unroll
                for (int K = N; K >= i; --K) {
                    const T c = recip<T>(K - i + 1);
                    const T apar =
                        (i == M + 1) ? ach[K] : astore[G::astore_offset(i - 1) + (K - i) * G::parents(i) + p];
                    const T apc = apar * c;
                    T acc = T(0);
#pragma unroll
                    for (int b = 0; b < W; ++b) {
                        T g;
                        if (K == i) {
                            g = Gt.word(i, s, b);  // the old cotangent of level i
                        } else {
                            g = src[(K - i - 1) * G::words(i) + p * W + b];
                            gsum[b] += g;
                        }
                        xbar[b] += g * apc;
                        acc += g * xr[b];
                    }
                    acc *= c;
                    if (i > M + 1) {
                        dst[(K - i) * G::parents(i) + p] = acc;
                    } else if (M == 1) {
                        // Abar^(K)_1 of the root word q: A_1 = S_0 x[q] / K + S_1[q]
                        Gt.root += acc;
                        xq += acc * S.s0 * recip<T>(K);
                        s0bar += acc * xs[q] * recip<T>(K);
                    } else {
                        // M = 0: Abar^(K)_0, the cotangent of S_0
                        s0bar += acc;
                    }
                }
#pragma unroll
                for (int b = 0; b < W; ++b) Gt.word(i, s, b) += gsum[b];
            }
        }
        __syncthreads();
    }
}

template <typename T, int W, int N, int NT, int M, int RL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Geo<W, N, NT, M, RL>::bwd_waves(sizeof(T)))))
k_sig_bwd_reg(DevBasis B, int64_t n_items, const int64_t* order, const T* incr, const int64_t* offs,
              View<const T> sigs, View<const T> cot, T* partial, int ck_every, const T* ckpts,
              const int64_t* ck_offs, unsigned* counter) {
    static_assert(M <= 1, "the backward pass keeps chain cotangents inside one work item");
    using G = Geo<W, N, NT, M, RL>;
    __shared__ T astore[G::astore_elems];
    __shared__ T lds0[G::lds_elems];
    __shared__ T lds1[G::lds_elems];
    __shared__ T sstate[G::state_elems];
    __shared__ T gstate[G::state_elems];
    __shared__ T xs[W];
    __shared__ T xneg[W];
    __shared__ T red[NT / 64][W];
    __shared__ unsigned slot;
    const int t = threadIdx.x;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) slot = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t item = slot;
        if (item >= n_items) break;
        const int64_t path = order ? order[item / G::P] : item / G::P;
        const int q = static_cast<int>(item % G::P);
        Tile<T, W, N, NT, M, RL> S(sstate), Gt(gstate);
        S.template move<false>(B, const_cast<T*>(sigs.obj(path)), sigs.cs, q, true);
        Gt.template move<false>(B, const_cast<T*>(cot.obj(path)), cot.cs, q, true);
        const int64_t t0 = offs[path], t1 = offs[path + 1];
        // increments are loaded one step ahead (see k_sig_fwd_reg)
        T xnext = (threadIdx.x < W && t0 < t1) ? incr[(t1 - 1) * W + threadIdx.x] : T(0);
        for (int64_t tt = t1 - 1; tt >= t0; --tt) {
            if (threadIdx.x < W) {
                xs[threadIdx.x] = xnext;
                xneg[threadIdx.x] = -xnext;
                if (tt - 1 >= t0) xnext = incr[(tt - 1) * W + threadIdx.x];
            }
            __syncthreads();
            const int64_t tl = tt - t0;
            if (tl == 0) {
                S.set_identity();
            } else if (ck_every > 0 && (tl % ck_every) == 0) {
                const T* src = ckpts + (ck_offs[path] + tl / ck_every - 1) * B.start[N + 1];
                S.template move<false>(B, const_cast<T*>(src), 1, q, true);
            } else {
                chen_step_reg<T, W, N, NT, M, RL>(S, xneg, lds0, lds1, q);
            }
            T xbar[W];
#pragma unroll
            for (int b = 0; b < W; ++b) xbar[b] = T(0);
            // Chain contribution to x[q] (M = 1), thread 0 only.  Target level
            // 1 first: new S_1[q] = S_0 x[q] + S_1[q], seeded by the old G_1[q]
            // before the higher targets add into it.
            T xq = (M == 1 && t == 0) ? Gt.root * S.s0 : T(0);
            T s0bar = T(0);  // unused: S_0 is the constant 1 along a path
            bwd_step<T, W, N, NT, M, RL>(S, Gt, xs, astore, lds0, lds1, q, xbar, xq, s0bar);
            // target level M = 1: root A_1 = S_0 x[q] + S_1[q], seeded by G_1[q]
            if (M == 1 && t == 0) {
#pragma unroll
                for (int b = 0; b < W; ++b)
                    if (b == q) xbar[b] += xq;
            }
            // work-group reduction of xbar (deterministic order)
#pragma unroll
            for (int b = 0; b < W; ++b) {
                T v = xbar[b];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
                if ((t & 63) == 0) red[t >> 6][b] = v;
            }
            __syncthreads();
            if (t < W) {
                T acc = T(0);
                for (int w = 0; w < NT / 64; ++w) acc += red[w][t];
                partial[(tt * G::P + q) * W + t] = acc;
            }
        }
    }
}

// Sum the per-prefix partials of k_fmexp_adj_tile in prefix order:
// grad_exponent[b] degree 1 (and 0 at degree 0 when its window holds it),
// and the S_0 cotangent added to grad_mult[b][0].
template <typename T, int W>
__global__ void k_scatter_partials(const T* partial, View<T> gx, View<T> gm, int64_t n, int P) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n * (W + 1)) return;
    const int64_t b = i / (W + 1);
    const int a = static_cast<int>(i - b * (W + 1));
    T acc = T(0);
    for (int q = 0; q < P; ++q) acc += partial[(b * P + q) * (W + 1) + a];
    if (a < W) {
        gx.obj(b)[(1 + a) * gx.cs] = acc;
    } else {
        gm.obj(b)[0] += acc;
        if (gx.mn == 0) gx.obj(b)[0] = T(0);
    }
}

// Per-operation adjoint of ft_fmexp with a degree-one exponent (one Chen
// step): given S = multiplier, x and G = d loss / d out, write
//   grad_mult = R_{exp(x)}^* G  and  grad_x = J(x)^T L_S^* G
// in one pass: the reverse sweep of k_sig_bwd_reg without the reconstruction.
// Each work item (object b, prefix q) writes W partial x sums and its share of
// the S_0 cotangent to partial[(b P + q)(W + 1) ..]; k_scatter_partials adds
// them up in a fixed order.
template <typename T, int W, int N, int NT, int M, int RL>
__global__ void __launch_bounds__(NT)
k_fmexp_adj_tile(DevBasis B, View<const T> mult, View<const T> ex, View<const T> cot, View<T> gmult, T* partial,
                 int64_t n) {
    static_assert(M <= 1, "chain cotangents stay inside one work item");
    using G = Geo<W, N, NT, M, RL>;
    __shared__ T astore[G::astore_elems];
    __shared__ T lds0[G::lds_elems];
    __shared__ T lds1[G::lds_elems];
    __shared__ T sstate[G::state_elems];
    __shared__ T gstate[G::state_elems];
    __shared__ T xs[W];
    __shared__ T red[NT / 64][W];
    const int t = threadIdx.x;
    for (int64_t item = blockIdx.x; item < n * G::P; item += gridDim.x) {
        const int64_t b = item / G::P;
        const int q = static_cast<int>(item - b * G::P);
        Tile<T, W, N, NT, M, RL> S(sstate), Gt(gstate);
        S.load_window(B, mult.obj(b), q, mult.mn, mult.mx);
        Gt.load_window(B, cot.obj(b), q, cot.mn, cot.mx);
        __syncthreads();
        if (t < W) xs[t] = ex.obj(b)[1 + t];
        __syncthreads();
        T xbar[W];
#pragma unroll
        for (int a = 0; a < W; ++a) xbar[a] = T(0);
        // target level 1 when M = 1: new S_1[q] = S_0 x[q] + S_1[q]
        T xq = (M == 1 && t == 0) ? Gt.root * S.s0 : T(0);
        T s0bar = (M == 1 && t == 0) ? Gt.root * xs[q] : T(0);
        bwd_step<T, W, N, NT, M, RL>(S, Gt, xs, astore, lds0, lds1, q, xbar, xq, s0bar);
        if (M == 1 && t == 0) {
#pragma unroll
            for (int a = 0; a < W; ++a)
                if (a == q) xbar[a] += xq;
        }
#pragma unroll
        for (int a = 0; a < W; ++a) {
            T v = xbar[a];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if ((t & 63) == 0) red[t >> 6][a] = v;
        }
        __syncthreads();
        if (t < W) {
            T acc = T(0);
            for (int w = 0; w < NT / 64; ++w) acc += red[w][t];
            partial[(b * G::P + q) * (W + 1) + t] = acc;
        }
        if (t == 0) partial[(b * G::P + q) * (W + 1) + W] = s0bar;
        // grad_mult: the cotangent of the old state; degree 0 holds G_0 here and
        // receives the summed S_0 cotangent from k_scatter_partials
        Gt.template move<true>(B, gmult.obj(b), 1, q, q == 0);
    }
}

template <typename T, int W, int N, int NT, int M, int RL>
int fmexp_adj_launch(const DevBasis& B, const rpp_hip_ft_fmexp_adjoint_desc& d, hipStream_t stream) {
    using G = Geo<W, N, NT, M, RL>;
    if constexpr (M > 1) {
        return fail(RPP_HIP_NOT_IMPLEMENTED, "adjoint tiles fix at most one letter");
    } else {
        const int64_t n = d.num_batches;
        void* part = nullptr;
        if (int rc = workspace(kSlotTmp, static_cast<size_t>(std::max<int64_t>(n, 1)) * G::P * (W + 1) * sizeof(T),
                               &part))
            return rc;
        const int64_t items = n * G::P;
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(items, 65535LL * 16));
        k_fmexp_adj_tile<T, W, N, NT, M, RL><<<grid, NT, 0, stream>>>(
            B, make_cview<T>(d.multiplier), make_cview<T>(d.exponent), make_cview<T>(d.cotangent),
            make_view<T>(d.grad_multiplier), static_cast<T*>(part), n);
        if (int rc = check_hip(hipGetLastError(), "k_fmexp_adj_tile launch")) return rc;
        k_scatter_partials<T, W><<<static_cast<unsigned>((n * (W + 1) + 255) / 256), 256, 0, stream>>>(
            static_cast<const T*>(part), make_view<T>(d.grad_exponent), make_view<T>(d.grad_multiplier), n, G::P);
        return check_hip(hipGetLastError(), "k_scatter_partials launch");
    }
}

__global__ void k_reset(unsigned* c) {
    if (threadIdx.x == 0) *c = 0u;
}

// Work-groups of `kernel` that fit on one compute unit at once (registers and
// LDS), at least 1; asked once per kernel.  A persistent grid of that many
// work-groups per compute unit keeps every resident slot busy.
template <typename K>
int resident_per_cu(K kernel, int threads) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess || blocks < 1)
        blocks = 1;
    return blocks;
}

// Partial rows are indexed by absolute increment row; rows outside
// [offsets[0], offsets[num_paths]) belong to no path, hold no partial sums and
// are left untouched in grad.
template <typename T, int W>
__global__ void k_sum_partials(const T* partial, T* grad, int64_t total, int P, const int64_t* offsets,
                               int64_t num_paths) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total * W) return;
    const int64_t t = i / W;
    if (t < offsets[0] || t >= offsets[num_paths]) return;
    const int b = static_cast<int>(i - t * W);
    T acc = T(0);
    for (int q = 0; q < P; ++q) acc += partial[(t * P + q) * W + b];
    grad[i] = acc;
}

template <typename T, int W, int N, int NT, int M, int RL>
int fwd_launch(const DevBasis& B, const rpp_hip_signature_desc& d, const int64_t* order, unsigned* counter,
               hipStream_t stream) {
    using G = Geo<W, N, NT, M, RL>;
    const int64_t items = d.num_paths * G::P;
    static const int per_cu = resident_per_cu(k_sig_fwd_reg<T, W, N, NT, M, RL>, NT);
    const unsigned grid =
        static_cast<unsigned>(std::min<int64_t>(items, static_cast<int64_t>(device_cu_count()) * per_cu));
    k_reset<<<1, 64, 0, stream>>>(counter);
    k_sig_fwd_reg<T, W, N, NT, M, RL><<<grid, NT, 0, stream>>>(
        B, items, order, static_cast<const T*>(d.increments), d.offsets, make_view<T>(d.out), d.checkpoint_every,
        static_cast<T*>(d.checkpoints), d.ckpt_offsets, counter);
    return check_hip(hipGetLastError(), "k_sig_fwd_reg launch");
}

template <typename T, int W, int N, int NT, int M, int RL>
int bwd_launch(const DevBasis& B, const rpp_hip_signature_backward_desc& d, const int64_t* order,
               int64_t total_incr, unsigned* counter, hipStream_t stream) {
    using G = Geo<W, N, NT, M, RL>;
    const int64_t items = d.num_paths * G::P;
    static const int per_cu = resident_per_cu(k_sig_bwd_reg<T, W, N, NT, M, RL>, NT);
    const unsigned grid =
        static_cast<unsigned>(std::min<int64_t>(items, static_cast<int64_t>(device_cu_count()) * per_cu));
    void* part = nullptr;
    if (int rc = workspace(kSlotTmp, static_cast<size_t>(std::max<int64_t>(total_incr, 1)) * G::P * W * sizeof(T), &part))
        return rc;
    k_reset<<<1, 64, 0, stream>>>(counter);
    k_sig_bwd_reg<T, W, N, NT, M, RL><<<grid, NT, 0, stream>>>(
        B, items, order, static_cast<const T*>(d.increments), d.offsets, make_cview<T>(d.signatures),
        make_cview<T>(d.cotangent), static_cast<T*>(part), d.checkpoint_every, static_cast<const T*>(d.checkpoints),
        d.ckpt_offsets, counter);
    if (int rc = check_hip(hipGetLastError(), "k_sig_bwd_reg launch")) return rc;
    if (total_incr > 0) {
        const int64_t n = total_incr * W;
        k_sum_partials<T, W><<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
            static_cast<const T*>(part), static_cast<T*>(d.grad_increments), total_incr, G::P, d.offsets,
            d.num_paths);
    }
    return check_hip(hipGetLastError(), "k_sum_partials launch");
}

}  // namespace

// The compiled tiles.  Each entry: dtype bytes, width, depth, threads,
// letters fixed per work-group, forward/backward launchers.
#define RPP_TILE(T, W, N, NT, M, RL)                                                               \
    {sizeof(T), W, N, NT, M, RL, &fwd_launch<T, W, N, NT, M, RL>,                                  \
     M <= 1 ? &bwd_launch<T, W, N, NT, (M <= 1 ? M : 1), RL> : nullptr,                             \
     M <= 1 ? &fmexp_adj_launch<T, W, N, NT, M, RL> : nullptr}

template <typename T>
const std::vector<RegTile<T>>& reg_tiles() {
    static const std::vector<RegTile<T>> tiles = {
        // (width, depth, threads, letters fixed, levels in registers); the
        // first tile of a shape is the library default
        // width 5, depth 6
        RPP_TILE(T, 5, 6, 512, 0, 6), RPP_TILE(T, 5, 6, 1024, 0, 6), RPP_TILE(T, 5, 6, 640, 1, 5),
        RPP_TILE(T, 5, 6, 320, 1, 5), RPP_TILE(T, 5, 6, 128, 2, 4), RPP_TILE(T, 5, 6, 320, 1, 1),
        RPP_TILE(T, 5, 6, 640, 1, 1), RPP_TILE(T, 5, 6, 320, 1, 2),
        // width 6, depth 6
        RPP_TILE(T, 6, 6, 448, 1, 5), RPP_TILE(T, 6, 6, 704, 1, 5), RPP_TILE(T, 6, 6, 1024, 1, 5),
        RPP_TILE(T, 6, 6, 256, 2, 4), RPP_TILE(T, 6, 6, 448, 1, 1), RPP_TILE(T, 6, 6, 704, 1, 1),
        RPP_TILE(T, 6, 6, 256, 1, 1), RPP_TILE(T, 6, 6, 448, 1, 2),
        // width 8, depth 5
        RPP_TILE(T, 8, 5, 256, 1, 4), RPP_TILE(T, 8, 5, 512, 1, 4), RPP_TILE(T, 8, 5, 64, 2, 3),
        RPP_TILE(T, 8, 5, 256, 1, 1), RPP_TILE(T, 8, 5, 512, 1, 1),
        // width 3, depth 8
        RPP_TILE(T, 3, 8, 768, 0, 8), RPP_TILE(T, 3, 8, 256, 1, 7), RPP_TILE(T, 3, 8, 512, 1, 7),
        RPP_TILE(T, 3, 8, 256, 2, 6), RPP_TILE(T, 3, 8, 256, 1, 1), RPP_TILE(T, 3, 8, 768, 1, 1),
        RPP_TILE(T, 3, 8, 256, 1, 2),
        // width 4, depth 4
        RPP_TILE(T, 4, 4, 64, 0, 4), RPP_TILE(T, 4, 4, 128, 0, 4), RPP_TILE(T, 4, 4, 64, 0, 1),
    };
    return tiles;
}

// The per-operation adjoint needs a tile fixing at most one letter; prefer
// one letter, then all levels in registers (one step per call leaves no
// recurrence for LDS-resident levels to pay off), then the smallest
// work-group.
template <typename T>
const RegTile<T>* fmexp_adj_tile(int width, int depth) {
    const RegTile<T>* best = nullptr;
    auto key = [](const RegTile<T>& t) { return std::make_tuple(t.split, t.reg_levels, -t.threads); };
    for (const auto& t : reg_tiles<T>()) {
        if (t.width != width || t.depth != depth || t.split > 1) continue;
        if (!best || key(t) > key(*best)) best = &t;
    }
    return best;
}

template const RegTile<float>* fmexp_adj_tile<float>(int, int);
template const RegTile<double>* fmexp_adj_tile<double>(int, int);
template const std::vector<RegTile<float>>& reg_tiles<float>();
template const std::vector<RegTile<double>>& reg_tiles<double>();

}  // namespace rpp_hip
