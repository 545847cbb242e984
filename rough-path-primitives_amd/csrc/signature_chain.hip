// Forward Chen recursion with two barriers per increment.
//
// The Horner scheme of one target level k,
//     A_0 = S_0,  A_i[u a] = A_{i-1}[u] x[a] / (k - i + 1) + S_i[u a],
//     new S_k = A_k,
// is a tree: A_i[w] depends only on the chain of prefixes of w.  Here a job
// (k, u) takes one word u of level k - D and
//   1. recomputes its chain A_{M..k-D}[prefixes of u] from the OLD lower
//      levels (held in LDS), k - D - M fused multiply-adds;
//   2. expands D levels below u in registers: W + W^2 + ... + W^D
//      multiply-adds, the last level being the new S_k[u v], |v| = D.
// Every job reads only old state, so all targets of one increment run at
// once; one barrier separates the reads of the old lower levels from their
// update, a second one publishes them.  The top level S_N stays in the
// registers of the jobs that own it (each job owns W^D words of S_N) for the
// whole path; levels M+1..N-1 live in LDS.
//
// A work item is (path, prefix q of M letters) as in signature_reg.hip: the
// words under q evolve independently given the chain S_0, S_1[q_1], ...,
// S_M[q], which every thread of the item carries in registers.
#include "common.hpp"
#include "signature.hpp"
#include "signature_chain.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int cpow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

template <int W, int N, int M, int D, int NT>
struct CGeo {
    static_assert(N >= M + 1 && D >= 1, "tile needs one level below the fixed prefix");
    static_assert(N - M >= D, "the top-level jobs expand D full levels");
    static constexpr int lvl(int i) { return cpow(W, i - M); }  // words of level i under q
    static constexpr int dexp(int k) { return (k - M) < D ? (k - M) : D; }
    static constexpr int jobs(int k) { return cpow(W, k - dexp(k) - M); }
    // jobs are numbered target N first, then N-1, ...
    static constexpr int job_begin(int k) {
        int b = 0;
        for (int kk = N; kk > k; --kk) b += jobs(kk);
        return b;
    }
    static constexpr int total_jobs = job_begin(M);
    static constexpr int slots = (total_jobs + NT - 1) / NT;
    static constexpr int R = cpow(W, D);
    static constexpr int low_off(int i) {
        int o = 0;
        for (int j = M + 1; j < i; ++j) o += lvl(j);
        return o;
    }
    static constexpr int low_size = low_off(N) > 0 ? low_off(N) : 1;
    static constexpr int P = cpow(W, M);
    static constexpr int T1 = D > 1 ? cpow(W, D - 1) : 1;  // expansion temporaries
};

template <typename T, int W, int N, int M, int D, int NT>
struct ChainState {
    using G = CGeo<W, N, M, D, NT>;
    T reg[G::slots][G::R];  // S_N words of top-level jobs, or the new S_k of a lower job
    T chain[M + 1];         // S_0, S_1[q_1], ..., S_M[q]
};

// Job of target K on ancestor a (level K - dexp(K), local index under q).
template <int K, typename T, int W, int N, int M, int D, int NT>
__device__ __forceinline__ void run_job(int a, T (&out)[CGeo<W, N, M, D, NT>::R], const T* chain_top,
                                        const T* low, const T* xs, const T* inv) {
    using G = CGeo<W, N, M, D, NT>;
    constexpr int De = G::dexp(K);
    constexpr int L0 = K - De;
    // chain down to level L0
    T A = chain_top[K];  // A_M of target K (uniform, precomputed)
#pragma unroll
    for (int i = M + 1; i <= L0; ++i) {
        const int anc = a / cpow(W, L0 - i);
        const int letter = anc % W;
        A = A * (xs[letter] * inv[K - i + 1]) + low[G::low_off(i) + anc];
    }
    // expand De levels: e = 1 .. De
    T cur[G::T1 > 0 ? G::T1 : 1];
    T nxt[G::T1 > 0 ? G::T1 : 1];
    cur[0] = A;
#pragma unroll
    for (int e = 1; e <= De; ++e) {
        constexpr int dummy = 0;
        (void)dummy;
        const int Le = L0 + e;
        const T c = inv[K - Le + 1];
        const int base = a * cpow(W, e);
#pragma unroll
        for (int r = 0; r < cpow(W, e); ++r) {
            const T v = cur[r / W] * (xs[r % W] * c);
            if (e == De) {
                if (Le == N) {
                    out[r] = v + out[r];
                } else {
                    out[r] = v + low[G::low_off(Le) + base + r];
                }
            } else {
                nxt[r] = v + low[G::low_off(Le) + base + r];
            }
        }
        if (e < De) {
#pragma unroll
            for (int r = 0; r < cpow(W, e); ++r) cur[r] = nxt[r];
        }
    }
}

template <int K, typename T, int W, int N, int M, int D, int NT>
__device__ __forceinline__ void dispatch_job(int j, T (&out)[CGeo<W, N, M, D, NT>::R], const T* chain_top,
                                             const T* low, const T* xs, const T* inv) {
    using G = CGeo<W, N, M, D, NT>;
    if constexpr (K >= M + 1) {
        if (j >= G::job_begin(K) && j < G::job_begin(K) + G::jobs(K)) {
            run_job<K, T, W, N, M, D, NT>(j - G::job_begin(K), out, chain_top, low, xs, inv);
        } else {
            dispatch_job<K - 1, T, W, N, M, D, NT>(j, out, chain_top, low, xs, inv);
        }
    }
}

// Write the new lower level held by a lower job into LDS.
template <int K, typename T, int W, int N, int M, int D, int NT>
__device__ __forceinline__ void publish_job(int j, const T (&vals)[CGeo<W, N, M, D, NT>::R], T* low) {
    using G = CGeo<W, N, M, D, NT>;
    if constexpr (K >= M + 1) {
        if (j >= G::job_begin(K) && j < G::job_begin(K) + G::jobs(K)) {
            constexpr int De = G::dexp(K);
            const int a = j - G::job_begin(K);
#pragma unroll
            for (int r = 0; r < cpow(W, De); ++r) low[G::low_off(K) + a * cpow(W, De) + r] = vals[r];
        } else {
            publish_job<K - 1, T, W, N, M, D, NT>(j, vals, low);
        }
    }
}

// Object layout helpers: global index of the word of level i with local
// index r under prefix q.
template <int W, int M>
__device__ __forceinline__ int64_t gidx(const DevBasis& B, int i, int64_t q, int64_t r) {
    return B.start[i] + q * static_cast<int64_t>(cpow(W, i - M)) + r;
}

template <typename T, int W, int N, int M, int D, int NT>
__device__ void store_state(const DevBasis& B, const ChainState<T, W, N, M, D, NT>& S, const T* low, T* o,
                            int64_t cs, int q) {
    using G = CGeo<W, N, M, D, NT>;
    const int t = threadIdx.x;
    // top level from the registers of the top-level jobs
#pragma unroll
    for (int s = 0; s < G::slots; ++s) {
        const int j = t + s * NT;
        if (j < G::jobs(N)) {
#pragma unroll
            for (int r = 0; r < G::R; ++r) o[gidx<W, M>(B, N, q, static_cast<int64_t>(j) * G::R + r) * cs] = S.reg[s][r];
        }
    }
    // levels M+1 .. N-1 from LDS
    for (int i = M + 1; i < N; ++i) {
        const int n = G::lvl(i);
        for (int r = t; r < n; r += NT) o[gidx<W, M>(B, i, q, r) * cs] = low[G::low_off(i) + r];
    }
    // the chain: level i < M word q / W^(M-i), written by the item whose
    // remaining letters are zero; level M by every item
    if (t == 0) {
#pragma unroll
        for (int i = 0; i <= M; ++i) {
            const int rest = cpow(W, M - i);
            if (q % rest == 0) o[(B.start[i] + q / rest) * cs] = S.chain[i];
        }
    }
}

template <typename T, int W, int N, int M, int D, int NT>
__global__ void __launch_bounds__(NT)
k_sig_fwd_chain(DevBasis B, int64_t n_items, const int64_t* order, const T* incr, const int64_t* offs, View<T> out,
                int ck_every, T* ckpts, const int64_t* ck_offs, unsigned* counter) {
    using G = CGeo<W, N, M, D, NT>;
    __shared__ T low[G::low_size];
    __shared__ T xs[2][W];
    __shared__ T inv[N + 2];
    __shared__ unsigned slot_sh;
    const int t = threadIdx.x;
    if (t <= N + 1) inv[t] = t == 0 ? T(1) : T(1) / T(t);
    const int64_t dim = B.start[N + 1];
    for (;;) {
        __syncthreads();
        if (t == 0) slot_sh = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t item = slot_sh;
        if (item >= n_items) break;
        const int64_t path = order ? order[item / G::P] : item / G::P;
        const int q = static_cast<int>(item % G::P);
        ChainState<T, W, N, M, D, NT> S;
#pragma unroll
        for (int s = 0; s < G::slots; ++s)
#pragma unroll
            for (int r = 0; r < G::R; ++r) S.reg[s][r] = T(0);
#pragma unroll
        for (int i = 0; i <= M; ++i) S.chain[i] = i == 0 ? T(1) : T(0);
        for (int r = t; r < G::low_size; r += NT) low[r] = T(0);
        const int64_t t0 = offs[path], t1 = offs[path + 1];
        if (t0 < t1 && t < W) xs[0][t] = incr[t0 * W + t];
        __syncthreads();
        for (int64_t tt = t0; tt < t1; ++tt) {
            const int buf = static_cast<int>((tt - t0) & 1);
            const T* x = xs[buf];
            if (tt + 1 < t1 && t < W) xs[buf ^ 1][t] = incr[(tt + 1) * W + t];
            // A_M of every target k > M along the prefix q (uniform)
            T chain_top[N + 1];
#pragma unroll
            for (int k = M + 1; k <= N; ++k) {
                T A = S.chain[0];
#pragma unroll
                for (int i = 1; i <= M; ++i) {
                    const int letter = (q / cpow(W, M - i)) % W;
                    A = A * (x[letter] * inv[k - i + 1]) + S.chain[i];
                }
                chain_top[k] = A;
            }
#pragma unroll
            for (int s = 0; s < G::slots; ++s) {
                const int j = t + s * NT;
                if (j < G::total_jobs) dispatch_job<N, T, W, N, M, D, NT>(j, S.reg[s], chain_top, low, x, inv);
            }
            __syncthreads();
            // the chain levels 1..M (targets k <= M), in descending order
#pragma unroll
            for (int k = M; k >= 1; --k) {
                T A = S.chain[0];
#pragma unroll
                for (int i = 1; i <= k; ++i) {
                    const int letter = (q / cpow(W, M - i)) % W;
                    A = A * (x[letter] * inv[k - i + 1]) + S.chain[i];
                }
                S.chain[k] = A;
            }
#pragma unroll
            for (int s = 0; s < G::slots; ++s) {
                const int j = t + s * NT;
                if (j >= G::jobs(N) && j < G::total_jobs) publish_job<N - 1, T, W, N, M, D, NT>(j, S.reg[s], low);
            }
            __syncthreads();
            if (ck_every > 0 && ((tt - t0 + 1) % ck_every) == 0) {
                T* dst = ckpts + (ck_offs[path] + (tt - t0 + 1) / ck_every - 1) * dim;
                store_state<T, W, N, M, D, NT>(B, S, low, dst, 1, q);
            }
        }
        store_state<T, W, N, M, D, NT>(B, S, low, out.obj(path), out.cs, q);
    }
}

__global__ void k_reset_chain(unsigned* c) {
    if (threadIdx.x == 0) *c = 0u;
}

template <typename T, int W, int N, int M, int D, int NT>
int chain_launch(const DevBasis& B, const rpp_hip_signature_desc& d, const int64_t* order, unsigned* counter,
                 hipStream_t stream) {
    using G = CGeo<W, N, M, D, NT>;
    const int64_t items = d.num_paths * G::P;
    const int64_t resident = static_cast<int64_t>(device_cu_count()) * std::max(1, 2048 / NT);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(items, resident)));
    k_reset_chain<<<1, 64, 0, stream>>>(counter);
    k_sig_fwd_chain<T, W, N, M, D, NT><<<grid, NT, 0, stream>>>(
        B, items, order, static_cast<const T*>(d.increments), d.offsets, make_view<T>(d.out), d.checkpoint_every,
        static_cast<T*>(d.checkpoints), d.ckpt_offsets, counter);
    return check_hip(hipGetLastError(), "k_sig_fwd_chain launch");
}

}  // namespace

#define RPP_CHAIN(T, W, N, M, D, NT) {W, N, M, D, NT, &chain_launch<T, W, N, M, D, NT>}

template <typename T>
const std::vector<ChainTile<T>>& chain_tiles() {
    static const std::vector<ChainTile<T>> tiles = {
        RPP_CHAIN(T, 5, 6, 1, 2, 192), RPP_CHAIN(T, 5, 6, 2, 2, 64), RPP_CHAIN(T, 5, 6, 0, 2, 832),
        RPP_CHAIN(T, 6, 6, 1, 2, 320), RPP_CHAIN(T, 6, 6, 2, 2, 64),
        RPP_CHAIN(T, 8, 5, 1, 2, 128), RPP_CHAIN(T, 8, 5, 2, 1, 64),
        RPP_CHAIN(T, 3, 8, 1, 3, 128), RPP_CHAIN(T, 3, 8, 2, 3, 64),
        RPP_CHAIN(T, 4, 4, 0, 2, 64),
    };
    return tiles;
}

template const std::vector<ChainTile<float>>& chain_tiles<float>();
template const std::vector<ChainTile<double>>& chain_tiles<double>();

}  // namespace rpp_hip
