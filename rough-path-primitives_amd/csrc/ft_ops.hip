// Per-operation kernels of the hip::block strategy: products (ft_mul, ft_fma,
// ft_inplace_fma, ft_inplace_mul), the composites (ft_exp, ft_fmexp, ft_log),
// tensor_pairing, the adjoint multiplications and sparse matrix-vector
// products.
//
// Execution model: one work-group per object of the batch (grid-stride over
// the batch), threads over the words of one level, output levels in
// DESCENDING degree order with a work-group barrier between levels.  An output
// word of degree k reads operand words of degree < k, plus at most the same
// word of degree k (the j = 0 or j = k split), which the same thread reads
// before it writes.  That is why out may alias an operand exactly as in the
// reference CPU strategy (ft_multiply.hpp:1596-1605).
//
// Word arithmetic never divides by a runtime level size: the split
// w -> (prefix, suffix) at j letters is advanced one letter at a time with a
// division by the width, compile-time for the common widths.
#include "common.hpp"
#include "ft_ops.hpp"
#include "fmexp_coal.hpp"
#include "log_lds.hpp"
#include "product_small.hpp"
#include "signature_reg.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int kBlock = 256;
// LDS budget of the staged small-tensor kernels (three 48 KiB work-groups
// per 160 KiB compute unit).
constexpr int64_t kLdsStageBytes = 48 * 1024;

template <bool CS1>
__device__ __forceinline__ int64_t at(int64_t idx, int64_t cs) {
    if constexpr (CS1) {
        return idx;
    } else {
        return idx * cs;
    }
}

// One pass of  out = [alpha * add] + beta * lhs (x) rhs  over out's window,
// degrees descending, each level finished by a barrier.  `add` is consulted
// only on levels inside its window (ScaledTileAddWriter semantics,
// ft_multiply.hpp:260-395); without it the level is beta * product
// (ScaledTileWriter, :176-258).  Each word's terms are summed in the
// reference's CPU order and every product and sum is rounded on its own
// (cpu_order_level_sum, contraction off), so the pass agrees with the
// reference bit for bit.  lhs[0] and rhs[0] are read before any level is
// written (degree 0 is the last one), so out may alias them.
template <typename T, int W, bool CS1>
__device__ void product_pass(const DevBasis& B, T* o, int64_t ocs, int omn, int omx,
                             const T* l, int64_t lcs, int lmn, int lmx,
                             const T* r, int64_t rcs, int rmn, int rmx,
                             const T* a, int64_t acs, int amn, int amx, bool with_add,
                             T alpha, T beta) {
#pragma clang fp contract(off)
    const bool lhs0_zero = lmn == 0 && lmx >= 0 && l[0] == T(0);
    const bool rhs0_zero = rmn == 0 && rmx >= 0 && r[0] == T(0);
    auto lval = [&](int d, uint32_t i) { return l[at<CS1>(B.start[d] + i, lcs)]; };
    auto rval = [&](int d, uint32_t i) { return r[at<CS1>(B.start[d] + i, rcs)]; };
    const int F = CpuTiling<W>::first(B), R = CpuTiling<W>::right(B);
    for (int k = omx; k >= omn; --k) {
        int jlo = max(rmn, k - lmx);
        int jhi = min(rmx, k - lmn);
        jlo = max(jlo, 0);
        jhi = min(jhi, k);
        const bool use_add = with_add && amn <= k && k <= amx;
        const uint32_t nk = B.pow[k];
        const int64_t ok = B.start[k];
        for (uint32_t w = threadIdx.x; w < nk; w += blockDim.x) {
            const T acc = cpu_order_level_sum<W, T>(B, F, R, k, w, jlo, jhi, lhs0_zero, rhs0_zero, lval, rval);
            T v = beta * acc;
            if (use_add) v = alpha * a[at<CS1>(ok + w, acs)] + v;
            o[at<CS1>(ok + w, ocs)] = v;
        }
        __syncthreads();
    }
}

template <typename T, bool CS1>
__device__ void fill_window(const DevBasis& B, T* o, int64_t cs, int mn, int mx, T v) {
    if (mn > mx) return;
    const int64_t b = B.start[mn], e = B.start[mx + 1];
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) o[at<CS1>(i, cs)] = v;
}

// ---------------------------------------------------------------- kernels

template <typename T, int W, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_product(DevBasis B, View<T> out, View<const T> lhs, View<const T> rhs, View<const T> add,
          int with_add, T alpha, T beta, int64_t n) {
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        product_pass<T, W, CS1>(B, out.obj(b), out.cs, out.mn, out.mx, lhs.obj(b), lhs.cs, lhs.mn,
                                lhs.mx, rhs.obj(b), rhs.cs, rhs.mn, rhs.mx,
                                with_add ? add.obj(b) : nullptr, add.cs, add.mn, add.mx,
                                with_add != 0, alpha, beta);
    }
}

// Division by a per-launch constant d >= 1 (n < 2^31): the multiply-shift of
// DevBasis::div_width, with the magic number formed on the device.
struct Divider {
    uint64_t magic;
    uint32_t shift;
    __device__ explicit Divider(uint32_t d) {
        uint32_t L = 0;
        while ((uint64_t{1} << L) < d) ++L;
        const uint64_t num = uint64_t{1} << (32 + L);
        magic = num / d + ((num % d) ? 1 : 0);
        shift = 32 + L;
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        return static_cast<uint32_t>((static_cast<uint64_t>(n) * magic) >> shift);
    }
};

// Small tensors: a work-group stages the operand windows of G objects in LDS
// with contiguous loads (object-major batches with stride = dim are one
// contiguous block), computes every output word of the G objects from LDS and
// writes each word once.  All operand reads precede the first write, so out
// may alias an operand; the addend is read at its own index.
template <typename T, int W>
__global__ void __launch_bounds__(kBlock)
k_product_lds(DevBasis B, View<T> out, View<const T> lhs, View<const T> rhs, View<const T> add,
              int with_add, T alpha, T beta, int64_t n, int G) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int F = CpuTiling<W>::first(B), R = CpuTiling<W>::right(B);
    T* sl = reinterpret_cast<T*>(smem_raw);
    const int64_t dim = B.start[B.depth + 1];
    T* sr = sl + G * dim;
    const bool lw = lhs.mn <= lhs.mx, rw = rhs.mn <= rhs.mx;
    const uint32_t lb = lw ? static_cast<uint32_t>(B.start[lhs.mn]) : 0;
    const uint32_t llen = lw ? static_cast<uint32_t>(B.start[lhs.mx + 1]) - lb : 0;
    const uint32_t rb = rw ? static_cast<uint32_t>(B.start[rhs.mn]) : 0;
    const uint32_t rlen = rw ? static_cast<uint32_t>(B.start[rhs.mx + 1]) - rb : 0;
    const uint32_t ob = static_cast<uint32_t>(B.start[out.mn]);
    const uint32_t nout = static_cast<uint32_t>(B.start[out.mx + 1]) - ob;
    const Divider dl(llen > 0 ? llen : 1), dr(rlen > 0 ? rlen : 1), dout(nout > 0 ? nout : 1);
    for (int64_t g0 = static_cast<int64_t>(blockIdx.x) * G; g0 < n; g0 += static_cast<int64_t>(gridDim.x) * G) {
        const uint32_t gn = static_cast<uint32_t>(min<int64_t>(G, n - g0));
        __syncthreads();
        for (uint32_t f = threadIdx.x; f < gn * llen; f += blockDim.x) {
            const uint32_t g = dl.div(f);
            const uint32_t i = lb + (f - g * llen);
            sl[g * dim + i] = lhs.obj(g0 + g)[i];
        }
        for (uint32_t f = threadIdx.x; f < gn * rlen; f += blockDim.x) {
            const uint32_t g = dr.div(f);
            const uint32_t i = rb + (f - g * rlen);
            sr[g * dim + i] = rhs.obj(g0 + g)[i];
        }
        __syncthreads();
        for (uint32_t f = threadIdx.x; f < gn * nout; f += blockDim.x) {
            const uint32_t g = dout.div(f);
            const uint32_t i = ob + (f - g * nout);
            int k = out.mn;
            while (i >= static_cast<uint32_t>(B.start[k + 1])) ++k;
            const uint32_t w = i - static_cast<uint32_t>(B.start[k]);
            const int jlo = max(max(rhs.mn, k - lhs.mx), 0);
            const int jhi = min(min(rhs.mx, k - lhs.mn), k);
            const T* Lg = sl + g * dim;
            const T* Rg = sr + g * dim;
            const bool lhs0_zero = lhs.mn == 0 && lw && Lg[0] == T(0);
            const bool rhs0_zero = rhs.mn == 0 && rw && Rg[0] == T(0);
            const T acc = cpu_order_level_sum<W, T>(
                B, F, R, k, w, jlo, jhi, lhs0_zero, rhs0_zero,
                [&](int d, uint32_t x) { return Lg[B.start[d] + x]; },
                [&](int d, uint32_t x) { return Rg[B.start[d] + x]; });
            T v = beta * acc;
            if (with_add && add.mn <= k && k <= add.mx) v = alpha * add.obj(g0 + g)[i] + v;
            out.obj(g0 + g)[i] = v;
        }
    }
}

// The composites in the reference's literal Horner order, all N steps in one
// launch (operations/intermediate/ft_exp.hpp:70-81, ft_fmexp.hpp:66-88,
// ft_log.hpp:70-84).  Used whenever the window layout leaves no shortcut.
template <typename T, int W, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_horner(int op, DevBasis B, View<T> out, View<const T> arg, View<const T> mult, int64_t n) {
    const int N = B.depth;
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T* o = out.obj(b);
        const T* x = arg.obj(b);
        if (op == kOpExp) {
            fill_window<T, CS1>(B, o, out.cs, out.mn, out.mx, T(0));
            __syncthreads();
            if (threadIdx.x == 0 && out.mn == 0 && out.mx >= 0) o[0] = T(1);
            __syncthreads();
            const int tmn = max(arg.mn, 1), tmx = min(arg.mx, N);
            for (int d = N; d > 0; --d) {
                const T divisor = T(1) / T(d);
                product_pass<T, W, CS1>(B, o, out.cs, out.mn, out.mx, o, out.cs, out.mn, out.mx, x,
                                        arg.cs, tmn, tmx, nullptr, 0, 1, 0, false, T(0), divisor);
                if (threadIdx.x == 0 && out.mn == 0 && out.mx >= 0) o[0] = o[0] + T(1);
                __syncthreads();
            }
        } else if (op == kOpFmexp) {
            const T* m = mult.obj(b);
            // VectorAssign (cpu/.../linalg/vector_assign.hpp:32-61)
            if (out.mn <= out.mx) {
                const int lo = max(out.mn, mult.mn), hi = min(out.mx, mult.mx);
                const int64_t bb = B.start[out.mn], ee = B.start[out.mx + 1];
                const int64_t cb = lo <= hi ? B.start[lo] : ee;
                const int64_t ce = lo <= hi ? B.start[hi + 1] : ee;
                for (int64_t i = bb + threadIdx.x; i < ee; i += blockDim.x)
                    o[at<CS1>(i, out.cs)] = (i >= cb && i < ce) ? m[at<CS1>(i, mult.cs)] : T(0);
            }
            __syncthreads();
            for (int d = N; d > 0; --d) {
                const int tmn = max(arg.mn, 1), tmx = min(arg.mx, N - d + 1);
                const T divisor = T(1) / T(d);
                product_pass<T, W, CS1>(B, o, out.cs, out.mn, out.mx, o, out.cs, out.mn, out.mx, x,
                                        arg.cs, tmn, tmx, m, mult.cs, mult.mn, mult.mx, true, T(1),
                                        divisor);
            }
        } else {  // kOpLog
            fill_window<T, CS1>(B, o, out.cs, out.mn, out.mx, T(0));
            __syncthreads();
            for (int d = N; d > 0; --d) {
                const T val = (d % 2 == 0 ? T(-1) : T(1)) / T(d);
                if (threadIdx.x == 0 && out.mn == 0 && out.mx >= 0) o[0] = o[0] + val;
                __syncthreads();
                const int omx = min(out.mx, N - d + 1);
                const int tmn = max(arg.mn, 1), tmx = min(arg.mx, N - d + 1);
                product_pass<T, W, CS1>(B, o, out.cs, out.mn, omx, o, out.cs, out.mn, omx, x, arg.cs,
                                        tmn, tmx, nullptr, 0, 1, 0, false, T(0), T(1));
            }
        }
        __syncthreads();
    }
}

// ft_fmexp when the exponent is a degree-one increment (window (1,1)) and
// out starts at degree 0, one word at a time: the reference's Horner
// recursion of the word's target level k along its prefixes,
//     R_0 = S_0 (+ 0 below the top level),  R_j = mult_j[w_1..w_j] + (0 + R_{j-1} x[w_j]) / (k-j+1),
// with the reference's roundings (horner_ref).  The prefixes come from
// repeated division (O(k^2) per word): the strided-layout fallback.
template <typename T, int W, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_fmexp_deg1(DevBasis B, View<T> out, View<const T> mult, View<const T> ex, int64_t n) {
    using L = Letters<W>;
    __shared__ T xs[256];
    __shared__ T inv[kMaxDepth + 2];
    const uint32_t width = L::width(B);
    const int N = B.depth;
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T* o = out.obj(b);
        const T* m = mult.obj(b);
        const T* x = ex.obj(b);
        for (uint32_t i = threadIdx.x; i < width; i += blockDim.x) xs[i] = x[at<CS1>(1 + i, ex.cs)];
        if (threadIdx.x <= static_cast<unsigned>(B.depth))
            inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
        __syncthreads();
        const T m0 = mult.has(0) ? m[0] : T(0);
        for (int k = out.mx; k >= 0; --k) {
            const uint32_t nk = B.pow[k];
            const int64_t ok = B.start[k];
            for (uint32_t w = threadIdx.x; w < nk; w += blockDim.x) {
                T r = k == N ? m0 : m0 + T(0);
                for (int j = 1; j <= k; ++j) {
                    uint32_t p = w;  // prefix of j letters: w / width^(k-j)
                    for (int e = j; e < k; ++e) p = L::div(B, p);
                    const uint32_t letter = p - L::div(B, p) * width;
                    const bool in = mult.has(j);
                    r = horner_ref(r, xs[letter], inv[k - j + 1], in ? m[at<CS1>(B.start[j] + p, mult.cs)] : T(0),
                                   in);
                }
                o[at<CS1>(ok + w, out.cs)] = r;
            }
        }
        __syncthreads();
    }
}

// ft_fmexp with a degree-one exponent x and out from degree 0, in Horner
// form per target level k (out must not alias the multiplier, R7):
//     A_0 = mult_0 (+ 0 below the top level),
//     A_i[p a] = mult_i[p a] + (0 + A_{i-1}[p] x[a]) / (k-i+1),   out_k = A_k,
// the reference's recursion with its roundings (horner_ref), at
// sum_k sum_{i<=k} d^i steps per object.
// A work item is (object, first letter q) when split = 1: the words under q
// evolve independently given the chain value A_1[q].  Thread t owns parents
// p = t, t + blockDim, ... of a level and computes their W children; the
// intermediates A_i (i < k) go through two LDS buffers.  mult and x are read
// from HBM once per item (lower levels again from cache), out written once.
template <typename T, int W>
__global__ void __launch_bounds__(kBlock)
k_fmexp_horner(DevBasis B, View<T> out, View<const T> mult, View<const T> ex, int64_t n, int split) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    const int M = split;
    const uint32_t P = M == 1 ? width : 1u;
    const uint32_t cap = N - 1 - M >= 0 ? B.pow[N - 1 - M] : 1u;
    T* buf0 = reinterpret_cast<T*>(smem_raw);
    T* buf1 = buf0 + cap;
    __shared__ T xs[256];
    __shared__ T inv[kMaxDepth + 2];
    if (threadIdx.x <= static_cast<unsigned>(N + 1))
        inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
    const int omx = out.mx;
    for (int64_t item = blockIdx.x; item < n * P; item += gridDim.x) {
        const int64_t b = item / P;
        const uint32_t q = static_cast<uint32_t>(item - b * P);
        T* o = out.obj(b);
        const T* m = mult.obj(b);
        const T* x = ex.obj(b);
        __syncthreads();
        for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) xs[a] = x[1 + a];
        __syncthreads();
        const T m0 = mult.has(0) ? m[0] : T(0);
        const bool in1 = mult.has(1);
        // chain value A_1[q] of target k (M = 1)
        const T m1q = (M == 1 && in1) ? m[1 + q] : T(0);
        if (threadIdx.x == 0) {
            if (q == 0 && omx >= 0) o[0] = N >= 1 ? m0 + T(0) : m0;
            if (M == 1 && omx >= 1) o[1 + q] = horner_ref(N == 1 ? m0 : m0 + T(0), xs[q], T(1), m1q, in1);
        }
        for (int k = omx; k >= M + 1; --k) {
            const T a0 = k == N ? m0 : m0 + T(0);
            const T chain = M == 1 ? horner_ref(a0, xs[q], inv[k], m1q, in1) : a0;
            for (int i = M + 1; i <= k; ++i) {
                const T c = inv[k - i + 1];
                const uint32_t np = B.pow[i - 1 - M];  // parents of level i in the item's subtree
                const int64_t base = B.start[i] + static_cast<int64_t>(q) * B.pow[i - M];
                const bool mi = mult.has(i);
                const T* prev = ((i - M) & 1) ? buf1 : buf0;
                T* cur = ((i - M) & 1) ? buf0 : buf1;
                for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
                    const T ap = i == M + 1 ? chain : prev[p];
                    const int64_t wb = base + static_cast<int64_t>(p) * width;
                    if constexpr (W > 0) {
#pragma unroll
                        for (int a = 0; a < W; ++a) {
                            const T v = horner_ref(ap, xs[a], c, mi ? m[wb + a] : T(0), mi);
                            if (i == k) o[wb + a] = v; else cur[p * W + a] = v;
                        }
                    } else {
                        for (uint32_t a = 0; a < width; ++a) {
                            const T v = horner_ref(ap, xs[a], c, mi ? m[wb + a] : T(0), mi);
                            if (i == k) o[wb + a] = v; else cur[p * width + a] = v;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// ft_exp of a degree-one argument, out from degree 0, one word at a time:
// the reference's Horner recursion (operations/intermediate/ft_exp.hpp:70-81)
// of the word's level k along its prefixes, R_0 = 1,
// R_j = (0 + R_{j-1} x[w_j]) / (k-j+1), with its roundings (horner_ref);
// exp(x)_w = prod x / k! up to rounding.  Prefixes by repeated division.
template <typename T, int W, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_exp_deg1(DevBasis B, View<T> out, View<const T> arg, int64_t n) {
    using L = Letters<W>;
    __shared__ T xs[256];
    __shared__ T inv[kMaxDepth + 2];
    const uint32_t width = L::width(B);
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T* o = out.obj(b);
        const T* x = arg.obj(b);
        for (uint32_t i = threadIdx.x; i < width; i += blockDim.x) xs[i] = x[at<CS1>(1 + i, arg.cs)];
        if (threadIdx.x <= static_cast<unsigned>(B.depth))
            inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
        __syncthreads();
        for (int k = 0; k <= out.mx; ++k) {
            const uint32_t nk = B.pow[k];
            const int64_t ok = B.start[k];
            for (uint32_t w = threadIdx.x; w < nk; w += blockDim.x) {
                T r = T(1);
                for (int j = 1; j <= k; ++j) {
                    uint32_t p = w;  // prefix of j letters
                    for (int e = j; e < k; ++e) p = L::div(B, p);
                    r = horner_ref(r, xs[p - L::div(B, p) * width], inv[k - j + 1], T(0), false);
                }
                o[at<CS1>(ok + w, out.cs)] = r;
            }
        }
        __syncthreads();
    }
}

// Work-group sum of one value per thread (wave64 shuffles, then LDS).
template <typename T>
__device__ T block_sum(T v) {
    __shared__ T part[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0)
        for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) r += part[i];
    __syncthreads();
    return r;
}

template <typename T, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_pairing(DevBasis B, T* out, int64_t os, View<const T> f, View<const T> x, int64_t n) {
    const int lo = max(f.mn, x.mn), hi = min(f.mx, x.mx);
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T acc = T(0);
        if (lo <= hi) {
            const T* F = f.obj(b);
            const T* X = x.obj(b);
            for (int64_t i = B.start[lo] + threadIdx.x; i < B.start[hi + 1]; i += blockDim.x)
                acc += F[at<CS1>(i, f.cs)] * X[at<CS1>(i, x.cs)];
        }
        acc = block_sum(acc);
        if (threadIdx.x == 0) out[b * os] = acc;
    }
}

// out[w] = sum_u op[u] arg[u w]   (left)   or   sum_u arg[w u] op[u]   (right),
// over op degrees pd with |w| + pd inside arg's window, pd descending as the
// reference's argument-degree-major loop (cpu/.../ft_adj_lmul.hpp:69-114),
// each product and sum rounded on its own: the reference's terms in its
// order, so both agree with it bit for bit (the right one through the
// antipode route of ft_adj_rmul.hpp:102-134).
template <typename T, bool LEFT, bool CS1>
__global__ void __launch_bounds__(kBlock)
k_adjoint(DevBasis B, View<T> out, View<const T> op, View<const T> arg, int64_t n, bool accumulate) {
#pragma clang fp contract(off)
    using L = Letters<0>;
    const uint32_t width = L::width(B);
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T* o = out.obj(b);
        const T* P = op.obj(b);
        const T* A = arg.obj(b);
        for (int od = out.mx; od >= out.mn; --od) {
            const int pmx = min(op.mx, arg.mx - od);
            const int pmn = max(op.mn, arg.mn - od);
            const uint32_t nk = B.pow[od];
            for (uint32_t w = threadIdx.x; w < nk; w += blockDim.x) {
                T acc = T(0);
                for (int pd = pmx; pd >= pmn; --pd) {
                    const int64_t ab = B.start[od + pd];
                    const int64_t pb = B.start[pd];
                    const uint32_t np = B.pow[pd];
                    if constexpr (LEFT) {
                        for (uint32_t u = 0; u < np; ++u)
                            acc += P[at<CS1>(pb + u, op.cs)] *
                                   A[at<CS1>(ab + static_cast<int64_t>(u) * nk + w, arg.cs)];
                    } else {
                        // the reference's route (antipodes around adj_lmul)
                        // takes the terms of op word v in the order of
                        // u = reverse(v) ascending: v follows u's letters
                        // with reversed weights, carried through u's
                        // trailing letters width - 1
                        const int64_t base = ab + static_cast<int64_t>(w) * np;
                        const uint32_t top = pd > 0 ? B.pow[pd - 1] : 0u;
                        uint32_t v = 0;
                        for (uint32_t u = 0; u < np; ++u) {
                            acc += P[at<CS1>(pb + v, op.cs)] * A[at<CS1>(base + v, arg.cs)];
                            uint32_t wgt = top, uu = u;
                            while (wgt > 0) {
                                const uint32_t q = L::div(B, uu);
                                if (uu - q * width != width - 1) break;
                                v -= (width - 1) * wgt;
                                wgt = L::div(B, wgt);
                                uu = q;
                            }
                            v += wgt;
                        }
                    }
                }
                T& dst = o[at<CS1>(B.start[od] + w, out.cs)];
                dst = accumulate ? dst + acc : acc;
            }
        }
    }
}

// tensor_antipode / tensor_reflect (cpu/.../tensor_generalised_antipode.hpp:
// 40-70): out[reverse(w)] = s_k arg[w] on each degree k of out's window, with
// s_k = (-1)^k when SIGN, 0 where arg lacks degree k.  Reversal is an
// involution, so each output word gathers arg[reverse(w)] and the writes stay
// coalesced.  One work-group per object.
template <typename T, int W, bool CS1, bool SIGN>
__global__ void __launch_bounds__(kBlock)
k_antipode(DevBasis B, View<T> out, View<const T> arg, int64_t n) {
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        T* o = out.obj(b);
        const T* a = arg.obj(b);
        for (int k = out.mn; k <= out.mx; ++k) {
            const bool has = arg.has(k);
            const T s = (SIGN && (k & 1)) ? T(-1) : T(1);
            const int64_t base = B.start[k];
            const uint32_t nk = B.pow[k];
            for (uint32_t w = threadIdx.x; w < nk; w += blockDim.x) {
                T v = T(0);
                if (has) {
                    uint32_t r = 0, x = w;
                    for (int d = 0; d < k; ++d) {
                        const uint32_t q = L::div(B, x);
                        r = r * width + (x - q * width);
                        x = q;
                    }
                    v = a[at<CS1>(base + r, arg.cs)] * s;
                }
                o[at<CS1>(base + w, out.cs)] = v;
            }
        }
    }
}

// CSR rows with the argument staged in LDS: a work-group streams objects; it
// copies the object's argument segment (the matrix columns) into LDS with
// coalesced loads, then each thread sums whole rows in entry order with
// contraction off, as the reference's loop does, so the results agree bit
// for bit.  The sparse matrix itself is shared by every object and stays in
// L2.  scale_arg: the reference's CSC product, whose scalar multiplies each
// argument value (rounded) before the products instead of the row sum
// (cpu/.../linalg/sparse_matrix_vector.hpp:150-158).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_spmv_csr_lds(T* out, int64_t obs, int64_t ocs, int64_t obeg, const T* arg, int64_t abs_, int64_t abeg,
               int64_t count, const T* values, const int64_t* indices, const int64_t* offsets, int64_t rows,
               T alpha, int scale_arg, int64_t n, int64_t cshift) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char spmv_raw[];
    T* xs = reinterpret_cast<T*>(spmv_raw);
    const int t = threadIdx.x;
    const T pre = scale_arg ? alpha : T(1);
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        const T* a = arg + b * abs_ + abeg;
        for (int64_t i = t; i < count; i += kBlock) xs[i] = scale_arg ? pre * a[i] : a[i];
        __syncthreads();
        T* o = out + b * obs;
        for (int64_t row = t; row < rows; row += kBlock) {
            const int64_t e0 = offsets[row], e1 = offsets[row + 1];
            T acc = T(0);
            for (int64_t e = e0; e < e1; ++e) acc += values[e] * xs[indices[e] - cshift];
            o[(obeg + row) * ocs] = scale_arg ? acc : alpha * acc;
        }
        __syncthreads();
    }
}

// CSR rows (cpu/.../linalg/sparse_matrix_vector.hpp:45-73): one thread per
// row, grid.y over objects.  The row sum runs in entry order with contraction
// off, as the reference's does, so the results agree bit for bit; scale_arg
// as for k_spmv_csr_lds.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_spmv_csr(T* out, int64_t obs, int64_t ocs, int64_t obeg, const T* arg, int64_t abs_,
           int64_t acs, int64_t abeg, const T* values, const int64_t* indices,
           const int64_t* offsets, int64_t rows, T alpha, int scale_arg, int64_t n, int64_t cshift) {
#pragma clang fp contract(off)
    const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const int64_t e0 = offsets[row], e1 = offsets[row + 1];
    for (int64_t b = blockIdx.y; b < n; b += gridDim.y) {
        const T* a = arg + b * abs_;
        T acc = T(0);
        if (scale_arg) {
            for (int64_t e = e0; e < e1; ++e) acc += values[e] * (alpha * a[(abeg + indices[e] - cshift) * acs]);
            out[b * obs + (obeg + row) * ocs] = acc;
        } else {
            for (int64_t e = e0; e < e1; ++e) acc += values[e] * a[(abeg + indices[e] - cshift) * acs];
            out[b * obs + (obeg + row) * ocs] = alpha * acc;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_vec_zero(T* out, int64_t obs, int64_t ocs, int64_t obeg, int64_t size, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size) return;
    for (int64_t b = blockIdx.y; b < n; b += gridDim.y) out[b * obs + (obeg + i) * ocs] = T(0);
}

bool all_contig(std::initializer_list<const rpp_hip_tensor_batch*> ts) {
    for (auto* t : ts)
        if (t && t->coeff_stride != 1) return false;
    return true;
}

// Width/contiguity dispatch: compile-time widths for contiguous batches.
#define RPP_DISPATCH_W(WIDTH, CONTIG, KERNEL_CALL)        \
    do {                                                 \
        if (CONTIG) {                                    \
            switch (WIDTH) {                             \
                case 2: { constexpr int W_ = 2; constexpr bool C_ = true; KERNEL_CALL; } break; \
                case 3: { constexpr int W_ = 3; constexpr bool C_ = true; KERNEL_CALL; } break; \
                case 4: { constexpr int W_ = 4; constexpr bool C_ = true; KERNEL_CALL; } break; \
                case 5: { constexpr int W_ = 5; constexpr bool C_ = true; KERNEL_CALL; } break; \
                case 6: { constexpr int W_ = 6; constexpr bool C_ = true; KERNEL_CALL; } break; \
                case 8: { constexpr int W_ = 8; constexpr bool C_ = true; KERNEL_CALL; } break; \
                default: { constexpr int W_ = 0; constexpr bool C_ = true; KERNEL_CALL; } break; \
            }                                            \
        } else {                                         \
            constexpr int W_ = 0; constexpr bool C_ = false; KERNEL_CALL; \
        }                                                \
    } while (0)

}  // namespace

// ------------------------------------------------------------ host launchers

template <typename T>
int launch_product(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& lhs, const rpp_hip_tensor_batch& rhs,
                   const rpp_hip_tensor_batch* add, double alpha, double beta, hipStream_t stream) {
    if (n <= 0) return 0;
    const View<T> o = make_view<T>(out);
    const View<const T> l = make_cview<T>(lhs), r = make_cview<T>(rhs);
    const View<const T> a = add ? make_cview<T>(*add) : View<const T>{nullptr, 0, 1, 1, 0};
    const int with_add = add ? 1 : 0;
    const bool contig = all_contig({&out, &lhs, &rhs, add});
    const int64_t dim = B.start[B.depth + 1];
    const int64_t per_object = 2 * dim * static_cast<int64_t>(sizeof(T));
    if (contig && out.min_degree <= out.max_degree && product_small_available<T>(B.width, B.depth))
        return launch_product_small<T>(B, n, out, lhs, rhs, add, alpha, beta, stream);
    if (contig && out.min_degree <= out.max_degree && per_object <= kLdsStageBytes) {
        const int G = static_cast<int>(std::min<int64_t>(64, kLdsStageBytes / per_object));
        const int64_t groups = (n + G - 1) / G;
        const unsigned grid = grid_for(groups);
        const size_t shmem = static_cast<size_t>(G) * per_object;
        RPP_DISPATCH_W(B.width, true,
                       (k_product_lds<T, W_><<<grid, kBlock, shmem, stream>>>(B, o, l, r, a, with_add, T(alpha),
                                                                             T(beta), n, G)));
        return check_hip(hipGetLastError(), "k_product_lds launch");
    }
    const unsigned grid = grid_for(n);
    RPP_DISPATCH_W(B.width, contig,
                   (k_product<T, W_, C_><<<grid, kBlock, 0, stream>>>(B, o, l, r, a, with_add, T(alpha),
                                                                      T(beta), n)));
    return check_hip(hipGetLastError(), "k_product launch");
}

template <typename T>
int launch_horner(int op, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                  const rpp_hip_tensor_batch& arg, const rpp_hip_tensor_batch* mult,
                  hipStream_t stream) {
    if (n <= 0) return 0;
    const View<T> o = make_view<T>(out);
    const View<const T> x = make_cview<T>(arg);
    const View<const T> m = mult ? make_cview<T>(*mult) : View<const T>{nullptr, 0, 1, 1, 0};
    const bool contig = all_contig({&out, &arg, mult});
    const unsigned grid = grid_for(n);
    // Shortcuts that compute the same truncated series in one pass.
    const bool deg1 = B.depth >= 1 && arg.max_degree == 1 && arg.min_degree <= 1 &&
                      out.min_degree == 0 && out.max_degree >= 0;
    if (op == kOpFmexp && deg1 && contig && out.max_degree == B.depth && mult && B.depth >= 1) {
        if (const FmexpTile<T>* tile = fmexp_coal_tile<T>(B.width, B.depth))
            return tile->launch(B, out, *mult, arg, n, stream);
    }
    if (op == kOpFmexp && deg1 && contig && B.width <= 256 && out.max_degree >= 1) {
        // split by first letter when the top-level exchange would not fit
        const int N = B.depth;
        const int split = (N >= 2 && static_cast<int64_t>(B.pow[N - 1]) * 2 * sizeof(T) > 16 * 1024) ? 1 : 0;
        const int64_t cap = N - 1 - split >= 0 ? B.pow[N - 1 - split] : 1;
        const size_t shmem = static_cast<size_t>(2 * cap) * sizeof(T);
        if (shmem <= static_cast<size_t>(kLdsStageBytes)) {
            const int64_t items = n * (split ? B.width : 1);
            const unsigned g = grid_for(items);
            RPP_DISPATCH_W(B.width, true,
                           (k_fmexp_horner<T, W_><<<g, kBlock, shmem, stream>>>(B, o, m, x, n, split)));
            return check_hip(hipGetLastError(), "k_fmexp_horner launch");
        }
    }
    if (op == kOpFmexp && deg1 && B.width <= 256) {
        RPP_DISPATCH_W(B.width, contig,
                       (k_fmexp_deg1<T, W_, C_><<<grid, kBlock, 0, stream>>>(B, o, m, x, n)));
        return check_hip(hipGetLastError(), "k_fmexp_deg1 launch");
    }
    if (op == kOpExp && deg1 && B.width <= 256) {
        RPP_DISPATCH_W(B.width, contig, (k_exp_deg1<T, W_, C_><<<grid, kBlock, 0, stream>>>(B, o, x, n)));
        return check_hip(hipGetLastError(), "k_exp_deg1 launch");
    }
    if (op == kOpLog && contig && out.min_degree == 0 && out.max_degree == B.depth &&
        log_lds_available<T>(B.width, B.depth))
        return launch_log_lds<T>(B, n, out, arg, stream);
    RPP_DISPATCH_W(B.width, contig,
                   (k_horner<T, W_, C_><<<grid, kBlock, 0, stream>>>(op, B, o, x, m, n)));
    return check_hip(hipGetLastError(), "k_horner launch");
}

template <typename T>
int launch_pairing(const DevBasis& B, int64_t n, const rpp_hip_scalar_batch& out,
                   const rpp_hip_tensor_batch& f, const rpp_hip_tensor_batch& x, hipStream_t stream) {
    if (n <= 0) return 0;
    const unsigned grid = grid_for(n);
    if (f.coeff_stride == 1 && x.coeff_stride == 1)
        k_pairing<T, true><<<grid, kBlock, 0, stream>>>(B, static_cast<T*>(out.data), out.stride,
                                                        make_cview<T>(f), make_cview<T>(x), n);
    else
        k_pairing<T, false><<<grid, kBlock, 0, stream>>>(B, static_cast<T*>(out.data), out.stride,
                                                         make_cview<T>(f), make_cview<T>(x), n);
    return check_hip(hipGetLastError(), "k_pairing launch");
}

template <typename T>
int launch_adjoint(bool left, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& op, const rpp_hip_tensor_batch& arg,
                   hipStream_t stream, bool accumulate) {
    if (n <= 0) return 0;
    const unsigned grid = grid_for(n);
    const bool contig = all_contig({&out, &op, &arg});
    const View<T> o = make_view<T>(out);
    const View<const T> p = make_cview<T>(op), g = make_cview<T>(arg);
    if (left) {
        if (contig) k_adjoint<T, true, true><<<grid, kBlock, 0, stream>>>(B, o, p, g, n, accumulate);
        else k_adjoint<T, true, false><<<grid, kBlock, 0, stream>>>(B, o, p, g, n, accumulate);
    } else {
        if (contig) k_adjoint<T, false, true><<<grid, kBlock, 0, stream>>>(B, o, p, g, n, accumulate);
        else k_adjoint<T, false, false><<<grid, kBlock, 0, stream>>>(B, o, p, g, n, accumulate);
    }
    return check_hip(hipGetLastError(), "k_adjoint launch");
}

template <typename T>
int launch_antipode(bool sign_by_degree, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                    const rpp_hip_tensor_batch& arg, hipStream_t stream) {
    if (n <= 0 || out.min_degree > out.max_degree) return 0;
    const unsigned grid = grid_for(n);
    const bool contig = all_contig({&out, &arg});
    const View<T> o = make_view<T>(out);
    const View<const T> a = make_cview<T>(arg);
    if (sign_by_degree) {
        RPP_DISPATCH_W(B.width, contig, (k_antipode<T, W_, C_, true><<<grid, kBlock, 0, stream>>>(B, o, a, n)));
    } else {
        RPP_DISPATCH_W(B.width, contig, (k_antipode<T, W_, C_, false><<<grid, kBlock, 0, stream>>>(B, o, a, n)));
    }
    return check_hip(hipGetLastError(), "k_antipode launch");
}

template <typename T>
int launch_spmv(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg,
                const rpp_hip_sparse_matrix& M, double alpha, hipStream_t stream, int64_t col_shift, bool csc) {
    if (n <= 0) return 0;
    if (M.format != RPP_HIP_CSR) return fail(RPP_HIP_INTERNAL, "launch_spmv takes CSR matrices");
    const unsigned gy = static_cast<unsigned>(std::min<int64_t>(n, 65535));
    T* o = static_cast<T*>(out.data);
    const T* a = static_cast<const T*>(arg.data);
    const T* v = static_cast<const T*>(M.values);
    const int scale_arg = csc ? 1 : 0;
    // CSC semantics zero the whole out view first: rows past the matrix's
    if (csc && out.size > M.rows) {
        dim3 zg(static_cast<unsigned>((out.size - M.rows + kBlock - 1) / kBlock), gy);
        k_vec_zero<T><<<zg, kBlock, 0, stream>>>(o, out.batch_stride, out.coeff_stride, out.begin + M.rows,
                                                 out.size - M.rows, n);
        if (int rc = check_hip(hipGetLastError(), "k_vec_zero launch")) return rc;
    }
    if (M.rows == 0) return 0;
    const int64_t count = std::min<int64_t>(M.cols, arg.size);
    const size_t lds = static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T);
    constexpr size_t kSpmvLdsMax = 128 * 1024;
    if (arg.coeff_stride == 1 && M.cols <= arg.size && lds <= kSpmvLdsMax && M.rows >= 64) {
        static const bool raised = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spmv_csr_lds<T>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       static_cast<int>(kSpmvLdsMax)) == hipSuccess;
        if (raised || lds <= 64 * 1024) {
            const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(8, (160 * 1024) / lds));
            const unsigned grid = static_cast<unsigned>(std::min<int64_t>(n, device_cu_count() * per_cu));
            k_spmv_csr_lds<T><<<grid, kBlock, lds, stream>>>(o, out.batch_stride, out.coeff_stride, out.begin, a,
                                                           arg.batch_stride, arg.begin, count, v, M.indices,
                                                           M.offsets, M.rows, T(alpha), scale_arg, n, col_shift);
            return check_hip(hipGetLastError(), "k_spmv_csr_lds launch");
        }
    }
    dim3 grid(static_cast<unsigned>((M.rows + kBlock - 1) / kBlock), gy);
    k_spmv_csr<T><<<grid, kBlock, 0, stream>>>(o, out.batch_stride, out.coeff_stride, out.begin, a, arg.batch_stride,
                                               arg.coeff_stride, arg.begin, v, M.indices, M.offsets, M.rows,
                                               T(alpha), scale_arg, n, col_shift);
    return check_hip(hipGetLastError(), "k_spmv_csr launch");
}

#define RPP_INSTANTIATE(T)                                                                      \
    template int launch_product<T>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,         \
                                   const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,      \
                                   const rpp_hip_tensor_batch*, double, double, hipStream_t);     \
    template int launch_horner<T>(int, const DevBasis&, int64_t, const rpp_hip_tensor_batch&,     \
                                  const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch*,       \
                                  hipStream_t);                                                   \
    template int launch_pairing<T>(const DevBasis&, int64_t, const rpp_hip_scalar_batch&,         \
                                   const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,      \
                                   hipStream_t);                                                  \
    template int launch_adjoint<T>(bool, const DevBasis&, int64_t, const rpp_hip_tensor_batch&,   \
                                   const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,      \
                                   hipStream_t, bool);                                            \
    template int launch_antipode<T>(bool, const DevBasis&, int64_t, const rpp_hip_tensor_batch&,  \
                                    const rpp_hip_tensor_batch&, hipStream_t);                     \
    template int launch_spmv<T>(int64_t, const rpp_hip_vector_batch&, const rpp_hip_vector_batch&, \
                                const rpp_hip_sparse_matrix&, double, hipStream_t, int64_t, bool);

RPP_INSTANTIATE(float)
RPP_INSTANTIATE(double)

}  // namespace rpp_hip
