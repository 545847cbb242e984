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

// Store a thread's subtree and the chain words it owns.  A wavefront's
// store instruction writes one word of each of 64 subtrees, so the lines
// fill in the L2 over the subtree's stores: non-temporal stores here measured
// several times slower (width 3 depth 8 over 16384 x 64 paths FP64, 1.36 ->
// 6.0 ms).
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

// Waves per SIMD a tile is compiled for (1: no target).  A SIMD holds 512
// registers per lane, so a tile a few registers above 512 / w runs w - 1
// waves.  The FP64 forward tiles at width 6 depth 6 (L = 4) and width 3
// depth 8 (L = 5) take 170 registers, two waves; held to 168 they run three
// with one 8-byte spill (2048 x 256 paths at (6,6): 2.93 -> 2.79 ms; 16384 x
// 64 paths at (3,8): 1.74 -> 1.69 ms).  The FP32 backward at width 6 depth 6
// (148 registers, three waves) held to 128 runs four (6.42 -> 6.08 ms).
// Larger cuts spill inside the step loop and run slower: the FP64 backward at
// width 6 depth 6 held to three waves, 11.6 -> 51 ms; the FP64 forward at
// width 5 depth 6 (L = 4) held to four, 1.38 -> 1.39 ms.  Results are the same
// bit for bit: only register allocation changes.
template <typename T, int W, int N, int L>
constexpr int thread_fwd_waves() {
    return sizeof(T) == 8 && ((W == 6 && N == 6 && L == 4) || (W == 3 && N == 8 && L == 5)) ? 3 : 1;
}
template <typename T, int W, int N, int L>
constexpr int thread_bwd_waves() {
    return sizeof(T) == 4 && W == 6 && N == 6 && L == 4 ? 4 : 1;
}

// CK: the forward also stores checkpoints.  Without them the step loop
// carries no checkpoint state or stores, which frees registers: FP64 width 5
// depth 6 (L = 4) 140 -> 108 (three waves per SIMD -> four), and the loop
// takes the form the compiler schedules best per type, measured over the
// benchmark shapes against the checkpointing loop (ms per launch): FP64 a
// ring of register sets indexed by step parity ((5,6) 4096 x 128 paths
// 1.23 -> 0.93, (6,6) 2048 x 256 2.55 -> 2.11, (3,8) 16384 x 64 1.47 -> 1.33,
// (8,5) 8192 uneven 9.47 -> 8.14; the plain loop without checkpoints 1.00,
// 2.30, 1.50, 8.78), FP32 the plain loop ((5,6) 0.61 -> 0.59, (6,6) 1.48 ->
// 1.33, (8,5) 5.39 -> 5.02, (3,8) 0.84 -> 0.80; the ring 0.66, 1.34, 5.61,
// 0.77).  The arithmetic is the same in every form, bit for bit.
// UL: the next increment's loads are issued with no test (the last step
// loads its own row again), so the compiler can wait for the current row
// alone instead of every load in flight.  Measured against the tested loads
// (ms per launch, alternating runs): FP64 width 6 depth 6 over 2048 x 256
// paths 2.51 -> 2.40, width 5 depth 6 over 4096 x 128 1.10 -> 1.07, width 3
// depth 8 over 16384 x 64 1.41 -> 1.385, FP32 width 8 depth 5 over the
// uneven 8192 paths 5.39 -> 5.29 (taken); FP64 width 8 depth 5 8.47 -> 8.70,
// FP32 width 5 depth 6 0.644 -> 0.72 (not taken).  The backward's loads
// measured no faster that way.
template <typename T, int W, int N, int L>
constexpr bool thread_fwd_uncond() {
    if constexpr (sizeof(T) == 8) {
        return (W == 5 && N == 6) || (W == 6 && N == 6) || (W == 3 && N == 8);
    } else {
        return W == 8 && N == 5;
    }
}

template <typename T, int W, int N, int L, int NT, bool CK>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(thread_fwd_waves<T, W, N, L>())))
k_sig_fwd_thread(DevBasis B, int n_paths, const int64_t* order, const T* __restrict__ incr, const int64_t* offs, View<T> out,
                 int ck_every, T* ckpts, const int64_t* ck_offs) {
    using G = ThreadGeo<W, N, L>;
    constexpr int LQ = L > 0 ? L : 1;
    constexpr bool UL = thread_fwd_uncond<T, W, N, L>();
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
    int ql[LQ];  // letters of q, first (most significant) first
#pragma unroll
    for (int i = 0; i < L; ++i) ql[i] = (q / tpow(W, L - 1 - i)) % W;
    T chain[L + 1];
    chain[0] = T(1);
#pragma unroll
    for (int i = 1; i <= L; ++i) chain[i] = T(0);
    T sub[G::total];
#pragma unroll
    for (int r = 0; r < G::total; ++r) sub[r] = T(0);
    // Increments one step ahead, in two register sets used in turn (the loop
    // is unrolled by two, so no copies): the row, uniform across the
    // wavefront, and the prefix letters' entries, loaded from the row pointer
    // with 32-bit lane offsets.
    auto load = [&](T (&x)[W], T (&xq)[LQ], int64_t t) {
        const T* row = incr + t * W;
#pragma unroll
        for (int b = 0; b < W; ++b) x[b] = row[b];
#pragma unroll
        for (int i = 0; i < L; ++i) xq[i] = row[ql[i]];
    };
    if constexpr (!CK && sizeof(T) == 8) {
        T xr[2][W], xqr[2][LQ];
        if (t0 < t1) load(xr[0], xqr[0], t0);
        for (int64_t t = t0; t < t1; t += 2) {
            static_for<2>([&](auto I) {
                constexpr int s = decltype(I)::value;
                const int64_t tt = t + s;
                if (tt < t1) {
                    if (UL || tt + 1 < t1) load(xr[s ^ 1], xqr[s ^ 1], tt + 1 < t1 ? tt + 1 : tt);
                    thread_step<T, W, N, L>(chain, sub, xr[s], xqr[s]);
                }
            });
        }
    } else {
        T xa[W], xqa[LQ], xb[W], xqb[LQ];
        // checkpoints: a countdown instead of a division per increment
        int ck_left = ck_every;
        int64_t ck_index = CK && ck_every > 0 ? ck_offs[path] : 0;
        const int64_t dim = B.start[N + 1];
        // a checkpoint holds the state below the top level, the part a
        // backward step reads (the top level is 83 % of the words at width 6
        // depth 6, so this writes a sixth of the bytes)
        auto checkpoint = [&]() {
            if (CK && ck_every > 0 && --ck_left == 0) {
                ck_left = ck_every;
                if (live) thread_store<T, W, N - 1, L, G::total>(B, ckpts + ck_index * dim, 1, opaque(q), chain, sub);
                ++ck_index;
            }
        };
        if (t0 < t1) load(xa, xqa, t0);
        for (int64_t t = t0; t < t1; t += 2) {
            if (UL || t + 1 < t1) load(xb, xqb, t + 1 < t1 ? t + 1 : t);
            thread_step<T, W, N, L>(chain, sub, xa, xqa);
            checkpoint();
            if (t + 1 >= t1) break;
            if (UL || t + 2 < t1) load(xa, xqa, t + 2 < t1 ? t + 2 : t + 1);
            thread_step<T, W, N, L>(chain, sub, xb, xqb);
            checkpoint();
        }
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
    const auto kernel = d.checkpoint_every > 0 ? &k_sig_fwd_thread<T, W, N, L, NT, true>
                                               : &k_sig_fwd_thread<T, W, N, L, NT, false>;
    kernel<<<grid, NT, 0, stream>>>(B, static_cast<int>(d.num_paths), order, static_cast<const T*>(d.increments),
                                    d.offsets, make_view<T>(d.out), d.checkpoint_every,
                                    static_cast<T*>(d.checkpoints), d.ckpt_offsets);
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
// Gbar[w_1..w_i] += Abar_i, Abar_{i-1} = Abar_i x[w_i] / (K-i+1).
//
// Levels >= L, thread-owned.  A thread rebuilds S_{t-1} below level N from
// S_t by a Chen step with -x (level N of S never enters the backward step),
// forms the Horner values along q and through its subtree, and walks the
// targets of its subtree and of its own chain word q down to level L:
//   - Gbar of its subtree words is complete in the thread (every extension of
//     a subtree word lies in the same subtree) and replaces G in place;
//   - Gbar of q (level L), the sum of those targets' Abar_L, is complete too;
//   - its share of xbar (subtree letters and q's last letter) is summed over
//     the work-group;
//   - Abar^(K)_{L-1}, one value per target level K, goes on to level L - 1.
// Levels < L, shared.  A chain word p of level i < L has W children of level
// i + 1; its walk sums their Abar^(K)_i, adds its own target (K = i, Abar =
// G[p]), and goes one level down.  Every such word has an executor thread,
// and level i's walk of a step runs L - i iterations after that step's
// thread-owned part, so the walks of all levels and the thread-owned work of
// later steps overlap, with one barrier per iteration handing each level its
// inputs.  The Horner values a walk needs are saved, when its step is
// computed, by the thread whose prefix continues p with zeros.  A step's xbar
// is complete after its level-1 walk and is then written, summed in a fixed
// order.  With the S_0 cotangent unused (S_0 is the constant 1 along a path)
// no value crosses work-groups.  Padding threads (r >= W^(L-1)) own no prefix
// and contribute nothing, but may execute walks.
template <int W, int N, int L>
struct BwdGeo {
    static constexpr int R = tpow(W, L - 1);        // threads with a prefix
    static constexpr int NT = (R + 63) / 64 * 64;  // padded to whole wavefronts
    static constexpr int NWV = NT / 64;
    static constexpr int count(int i) { return tpow(W, i - 1); }  // chain words of level i under a
    // level i's walk executors: threads [exec(i), exec(i) + count(i)); levels
    // L-1 .. 1 in turn, each from a new wavefront while that fits
    static constexpr int exec(int i) {
        int b = 0;
        for (int j = L - 1; j > i; --j) {
            const int end = b + count(j);
            const int up = (end + 63) / 64 * 64;
            b = (up + count(j - 1) <= NT) ? up : end;
        }
        return b;
    }
    // inputs of level i's walks: Abar^(K)_i, K = i+1..N, of each level-(i+1) word
    static constexpr int vin(int i) { return N - i; }
    static constexpr int voff(int i) {
        int o = 0;
        for (int j = 1; j < i; ++j) o += count(j + 1) * vin(j);
        return o;
    }
    static constexpr int vtotal = voff(L) > 0 ? voff(L) : 1;
    // saved for level i's walk of a step: A^(K)_{i-1}, K = i..N, and x[p_i]
    static constexpr int cin(int i) { return N - i + 2; }
    static constexpr int coff(int i) {
        int o = 0;
        for (int j = 1; j < i; ++j) o += count(j) * cin(j);
        return o;
    }
    static constexpr int ctotal = coff(L) > 0 ? coff(L) : 1;
    // a step's xbar share from each chain word of levels < L
    static constexpr int xoff(int i) {
        int o = 0;
        for (int j = 1; j < i; ++j) o += count(j);
        return o;
    }
    static constexpr int xtotal = xoff(L) > 0 ? xoff(L) : 1;
};

// Cross-lane moves of a 32- or 64-bit value: the two halves of a 64-bit one
// move separately.
template <typename T, typename F>
__device__ __forceinline__ T lanes_move(T v, F&& f) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = f(static_cast<uint32_t>(u)), hi = f(static_cast<uint32_t>(u >> 32));
        return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
    } else {
        return __builtin_bit_cast(T, f(__builtin_bit_cast(uint32_t, v)));
    }
}

// The value of the lane that the DPP control CTRL names (within a row of 16
// lanes).  Every control used here names a lane for every lane, so the move
// needs no prior value in its destination.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_move(uint32_t u) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(u), CTRL, 0xf, 0xf, true));
}

// x + the value of the lane that the DPP control CTRL names.
template <int CTRL, typename T>
__device__ __forceinline__ T add_dpp(T x) {
    return x + lanes_move(x, [](uint32_t u) { return dpp_move<CTRL>(u); });
}

// x summed over the lane pairs that differ in lane bit HALF (32 or 16) in
// the lanes where that bit is clear, y summed over them where it is set.
// 64-bit values: the two results of v_permlane32_swap / v_permlane16_swap
// (x of the lower half of the lanes with y of the lower half, x of the upper
// half with y of the upper half; even and odd rows of 16 for HALF = 16),
// added.  32-bit values: a select and a lane shuffle, since the compiler here
// drops the second result of a 32-bit swap that feeds a float add.
template <int HALF, typename T>
__device__ __forceinline__ T swap_sum(T x, T y, int lane) {
    if constexpr (sizeof(T) == 8) {
        auto sw = [](uint32_t a, uint32_t b) {
            if constexpr (HALF == 32) {
                return __builtin_amdgcn_permlane32_swap(a, b, false, false);
            } else {
                return __builtin_amdgcn_permlane16_swap(a, b, false, false);
            }
        };
        const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
        const auto lo = sw(static_cast<uint32_t>(ux), static_cast<uint32_t>(uy));
        const auto hi = sw(static_cast<uint32_t>(ux >> 32), static_cast<uint32_t>(uy >> 32));
        const T a = __builtin_bit_cast(T, (static_cast<uint64_t>(hi[0]) << 32) | lo[0]);
        const T b = __builtin_bit_cast(T, (static_cast<uint64_t>(hi[1]) << 32) | lo[1]);
        return a + b;
    } else {
        const bool up = (lane & HALF) != 0;
        const T send = up ? x : y;
        const T keep = up ? y : x;
        return keep + __shfl_xor(send, HALF, 64);
    }
}

// Sums W <= 8 per-lane values over the 64 lanes of a wavefront.  The values
// sit in P = 2^LP >= W slots; in each of the first LP exchanges the lanes that
// differ in one lane bit (32, 16, 8) trade the half of their slots the other
// lane keeps, so each lane goes on with half as many partial sums: the 32-
// and 16-lane exchanges by lane swaps, the 8-lane one by a DPP rotation.  DPP
// butterflies over the remaining lane bits finish the sums.  Lane l then
// holds the total of slot l >> (6 - LP).  A fixed order, so the sums are
// deterministic.
template <typename T, int W>
struct WaveSlots {
    static_assert(W >= 1 && W <= 8, "at most 8 slots");
    static constexpr int LP = W <= 1 ? 0 : W <= 2 ? 1 : W <= 4 ? 2 : 3;
    static constexpr int P = 1 << LP;
    __device__ static __forceinline__ T sum(const T (&v)[W], int lane) {
        T s[P];
#pragma unroll
        for (int b = 0; b < P; ++b) s[b] = b < W ? v[b] : T(0);
        if constexpr (LP >= 1) {
#pragma unroll
            for (int j = 0; j < P / 2; ++j) s[j] = swap_sum<32>(s[j], s[j + P / 2], lane);
        } else {
            s[0] = swap_sum<32>(s[0], s[0], lane);  // both halves: each lane's pair sum
        }
        if constexpr (LP >= 2) {
#pragma unroll
            for (int j = 0; j < P / 4; ++j) s[j] = swap_sum<16>(s[j], s[j + P / 4], lane);
        } else {
            s[0] = swap_sum<16>(s[0], s[0], lane);
        }
        if constexpr (LP >= 3) {
            const bool up = (lane & 8) != 0;
            const T send = up ? s[0] : s[1];
            const T keep = up ? s[1] : s[0];
            s[0] = keep + lanes_move(send, [](uint32_t u) { return dpp_move<0x128>(u); });  // row_ror:8
        } else {
            s[0] = add_dpp<0x128>(s[0]);  // row_ror:8
        }
        s[0] = add_dpp<0x141>(s[0]);  // row_half_mirror
        s[0] = add_dpp<0x4E>(s[0]);   // quad_perm [2,3,0,1]
        s[0] = add_dpp<0xB1>(s[0]);   // quad_perm [1,0,3,2]
        return s[0];
    }
    __device__ static __forceinline__ int slot(int lane) { return lane >> (6 - LP); }
    __device__ static __forceinline__ bool leader(int lane) { return (lane & ((1 << (6 - LP)) - 1)) == 0; }
};

// CK: S_{t-1} is reloaded from the forward's checkpoints where one was
// stored.  The instance without checkpoints has no reload path in its step
// loop (width 6 depth 6 over 2048 x 256 paths: FP64 11.67 -> 11.41 ms, FP32
// 6.13 -> 5.47 ms; width 5 depth 6 FP64 5.22 -> 5.03 ms), bit-identical.
template <typename T, int W, int N, int L, bool CK, int NT = BwdGeo<W, N, L>::NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(thread_bwd_waves<T, W, N, L>())))
k_sig_bwd_thread(DevBasis B, int64_t n_items, const int64_t* order, const T* incr, const int64_t* offs,
                 View<const T> sigs, View<const T> cot, T* partial, int ck_every, const T* ckpts,
                 const int64_t* ck_offs) {
    static_assert(L >= 1 && L <= N - 2, "the prefix leaves a subtree below level N");
    using GS = ThreadGeo<W, N - 1, L>;  // S below level N
    using GG = ThreadGeo<W, N, L>;      // G, and the Horner values
    using BG = BwdGeo<W, N, L>;
    using WS = WaveSlots<T, W>;
    constexpr int R = BG::R;
    constexpr int NWV = BG::NWV;
    static_assert(L == 1 || BG::exec(1) + 1 <= NT, "walk executors fit in the work-group");
    __shared__ T vbuf[2][BG::vtotal];        // walk inputs, by iteration parity
    __shared__ T cbuf[L][BG::ctotal];        // saved Horner values, by step mod L
    __shared__ T xsb[L + 1][BG::xtotal];     // walks' xbar shares, by step mod (L + 1)
    __shared__ T red[L + 1][NWV][W];         // wavefront xbar sums, by step mod (L + 1)
    const int64_t item = blockIdx.x;
    if (item >= n_items) return;
    const int64_t slot = item / W;
    const int a = static_cast<int>(item - slot * W);
    const int64_t path = order ? order[slot] : slot;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const bool live = t < R;
    const int q = a * R + (live ? t : 0);
    int ql[L];
#pragma unroll
    for (int i = 0; i < L; ++i) ql[i] = (q / tpow(W, L - 1 - i)) % W;
    // S_t (below level N) and G_t: the thread's subtree, its chain copy and
    // the cotangent of its chain word q
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
    T gL = g0[(B.start[L] + q) * cot.cs];
    // Padding threads hold zero cotangents, so every value they add to the
    // xbar sums is zero and no step needs to mask them out.
    if (!live) {
        gL = T(0);
#pragma unroll
        for (int rr = 0; rr < GG::total; ++rr) subG[rr] = T(0);
    }
    // the walk this thread executes, if any: level ex, chain word ej under a,
    // and that word's cotangent
    int ex = 0, ej = 0;
#pragma unroll
    for (int i = 1; i < L; ++i)
        if (t >= BG::exec(i) && t < BG::exec(i) + BG::count(i)) {
            ex = i;
            ej = t - BG::exec(i);
        }
    T gst = T(0);
    if (ex > 0) gst = g0[(B.start[ex] + static_cast<int64_t>(a) * tpow(W, ex - 1) + ej) * cot.cs];
    const int64_t t0 = offs[path], t1 = offs[path + 1];
    const int64_t len = t1 - t0;
    const int64_t dim = B.start[N + 1];
    // increments one step ahead: the row (uniform) and the prefix letters'
    T xn[W], xqn[L];
    if (len > 0) {
#pragma unroll
        for (int b = 0; b < W; ++b) xn[b] = incr[(t1 - 1) * W + b];
        const T* row = incr + (t1 - 1) * W;
#pragma unroll
        for (int i = 0; i < L; ++i) xqn[i] = row[ql[i]];
    }
    // iteration r: the thread-owned part of step t1-1-r (r < len), level i's
    // walks of step t1-1-(r-(L-i)), and the xbar of step t1-1-(r-L)
    // ring slots of iteration r: r mod 2, r mod L and r mod (L + 1), kept as
    // counters
    const int64_t iters = len > 0 ? len + L : 0;
    int cslot = 0, xslot = 0;
    for (int64_t r = 0; r < iters; ++r, cslot = (cslot + 1 == L ? 0 : cslot + 1),
                 xslot = (xslot == L ? 0 : xslot + 1)) {
        const int vslot = static_cast<int>(r & 1);
        if (r < len) {
            const int64_t tt = t1 - 1 - r;
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
                const T* row = incr + (tt - 1) * W;
#pragma unroll
                for (int i = 0; i < L; ++i) xqn[i] = row[ql[i]];
            }
            // S_{t-1}: identity before the first increment, a checkpoint
            // where one was stored, else S_t (x) exp(-x)
            const int64_t tl = tt - t0;
            if (tl == 0) {
                chain[0] = T(1);
#pragma unroll
                for (int i = 1; i <= L; ++i) chain[i] = T(0);
#pragma unroll
                for (int rr = 0; rr < GS::total; ++rr) subS[rr] = T(0);
            } else if (CK && ck_every > 0 && (tl % ck_every) == 0) {
                const T* c0 = ckpts + (ck_offs[path] + tl / ck_every - 1) * dim;
                const int qo = opaque(q);
#pragma unroll
                for (int i = 0; i <= L; ++i) chain[i] = c0[B.start[i] + qo / tpow(W, L - i)];
#pragma unroll
                for (int i = L + 1; i < N; ++i)
#pragma unroll
                    for (int rr = 0; rr < GS::words(i); ++rr)
                        subS[GS::offset(i) + rr] = c0[B.start[i] + static_cast<int64_t>(qo) * GS::words(i) + rr];
            } else {
                thread_step<T, W, N - 1, L>(chain, subS, xneg, xqneg);
            }
            // Horner values of the step S_{t-1} -> S_t along q: hz[K][i] =
            // A^(K)_i for i < min(K, L), ach[K] = A^(K)_L for K > L
            T hz[N + 1][L];
            T ach[N + 1];
#pragma unroll
            for (int K = 1; K <= N; ++K) {
                T v = chain[0];
                hz[K][0] = v;
#pragma unroll
                for (int i = 1; i < L; ++i)
                    if (i < K) {
                        v = (v * trecip<T>(K - i + 1)) * xq[i - 1] + chain[i];
                        hz[K][i] = v;
                    }
                if (K > L) ach[K] = (v * trecip<T>(K - L + 1)) * xq[L - 1] + chain[L];
            }
            // the Horner values of the walks of levels < L, saved by the
            // thread whose prefix continues the chain word with zeros
#pragma unroll
            for (int i = 1; i < L; ++i)
                if (live && t % tpow(W, L - i) == 0) {
                    T* c = &cbuf[cslot][BG::coff(i) + (t / tpow(W, L - i)) * BG::cin(i)];
#pragma unroll
                    for (int K = i; K <= N; ++K) c[K - i] = hz[K][i - 1];
                    c[N - i + 1] = xq[i - 1];
                }
            // Horner values of subtree levels i < N
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
            // reverse over the subtree, levels descending: ab holds Abar^(K)_i
            // of the level below the one processed (targets K > i), abl
            // Abar^(K)_L
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
                            const T prev =
                                (i == L + 1) ? ach[K] : av[GG::aoffset(i - 1) + (K - i) * GG::words(i - 1) + u];
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
            // level L: the subtree targets and q's own, on to level L - 1
            T cl = T(0), xl = T(0);
            T vout[N - L + 1];
#pragma unroll
            for (int K = N; K >= L; --K) {
                const T abk = (K == L) ? gL : abl[K];
                cl += abk;
                const T abc = abk * trecip<T>(K - L + 1);
                xl += abc * hz[K][L - 1];
                vout[K - L] = abc * xq[L - 1];
            }
            gL = cl;
#pragma unroll
            for (int b = 0; b < W; ++b)
                if (ql[L - 1] == b) xbar[b] += xl;
            if (live && L > 1) {
#pragma unroll
                for (int k = 0; k <= N - L; ++k) vbuf[vslot][BG::voff(L - 1) + t * BG::vin(L - 1) + k] = vout[k];
            }
            const T xs = WS::sum(xbar, lane);
            if (WS::leader(lane) && WS::slot(lane) < W) red[xslot][t >> 6][WS::slot(lane)] = xs;
        }
        // the walks of levels < L
        static_for<(L > 1 ? L - 1 : 0)>([&](auto I) {
            constexpr int i = decltype(I)::value + 1;
            if (ex != i) return;
            // the step of this walk was computed L - i iterations ago
            const int64_t rs = r - (L - i);
            if (rs < 0 || rs >= len) return;
            const int cs = cslot >= L - i ? cslot - (L - i) : cslot + i;
            const int xs = xslot >= L - i ? xslot - (L - i) : xslot + i + 1;
            const T* vi = &vbuf[vslot ^ 1][BG::voff(i) + ej * W * BG::vin(i)];
            T V[N - i];
#pragma unroll
            for (int k = 0; k < N - i; ++k) {
                T acc = T(0);
#pragma unroll
                for (int c = 0; c < W; ++c) acc += vi[c * BG::vin(i) + k];
                V[k] = acc;
            }
            const T* cf = &cbuf[cs][BG::coff(i) + ej * BG::cin(i)];
            T ci = T(0), xw = T(0);
#pragma unroll
            for (int K = N; K >= i; --K) {
                const T abk = (K == i) ? gst : V[K - i - 1];
                ci += abk;
                const T abc = abk * trecip<T>(K - i + 1);
                xw += abc * cf[K - i];
                if (i > 1) vbuf[vslot][BG::voff(i - 1) + ej * BG::vin(i - 1) + (K - i)] = abc * cf[N - i + 1];
            }
            gst = ci;
            xsb[xs][BG::xoff(i) + ej] = xw;
        });
        // a step whose walks are all done: its xbar, in a fixed order
        const int64_t rf = r - L;
        if (rf >= 0 && rf < len && t < W) {
            const int sl = xslot == L ? 0 : xslot + 1;  // (r - L) mod (L + 1)
            T acc = T(0);
#pragma unroll
            for (int w = 0; w < NWV; ++w) acc += red[sl][w][t];
            static_for<(L > 1 ? L - 1 : 0)>([&](auto I) {
                constexpr int i = L - 1 - decltype(I)::value;  // levels L-1 .. 1
                if constexpr (i == 1) {
                    if (a == t) acc += xsb[sl][BG::xoff(1)];
                } else {
#pragma unroll
                    for (int j = 0; j < BG::count(i) / W; ++j) acc += xsb[sl][BG::xoff(i) + j * W + t];
                }
            });
            partial[((t1 - 1 - rf) * W + a) * W + t] = acc;
        }
        __syncthreads();
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
    const auto kernel = d.checkpoint_every > 0 ? &k_sig_bwd_thread<T, W, N, L, true>
                                               : &k_sig_bwd_thread<T, W, N, L, false>;
    kernel<<<static_cast<unsigned>(items), BwdGeo<W, N, L>::NT, 0, stream>>>(
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
