// Batched Chen recursion over paths of uneven length: forward signatures and
// their backward pass, one persistent launch each.
//
// One Chen step S <- S (x) exp(x) for a degree-one increment x is evaluated
// level by level with a Horner scheme per target level k (descending, in
// place):
//     A_0 = S_0,   A_i[p a] = A_{i-1}[p] * x[a] / (k - i + 1) + S_i[p a],
//     new S_k = A_k,
// which is sum_i S_i (x) x^{k-i}/(k-i)!  -- the truncated product the
// reference's ft_fmexp computes with N full passes
// (operations/intermediate/ft_fmexp.hpp:66-88), here at
// sum_k sum_{i<=k} d^i multiply-adds per increment.
//
// The backward pass walks the increments in reverse.  It reconstructs the
// previous state with S_{t-1} = S_t (x) exp(-x_t) (or reloads it from a
// forward checkpoint), then runs reverse-mode through the same Horner
// scheme: for each target level k (ascending, so the cotangent can be
// updated in place),
//     Abar_k = G_k;  Abar_{i-1}[p] = sum_a Abar_i[p a] x[a] c,
//     xbar[a] += sum_p Abar_i[p a] A_{i-1}[p] c,   G_{i-1} += Abar_{i-1},
// with c = 1/(k - i + 1) and A_i the forward Horner values.
//
// This file holds the storage-generic kernels: any (width <= 16 for the
// backward, depth), state and Horner intermediates in a per-work-group slice
// of a device workspace.  Paths are taken from a length-sorted work queue.
#include "common.hpp"
#include "signature.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace rpp_hip {

namespace {

constexpr int kSigBlock = 256;
constexpr int kMaxBackwardWidth = 16;

__global__ void k_reset_counter(unsigned* c) {
    if (threadIdx.x == 0) *c = 0u;
}

// Take the next queue position for this work-group (broadcast through LDS).
__device__ __forceinline__ int64_t next_item(unsigned* counter) {
    __shared__ unsigned slot;
    __syncthreads();
    if (threadIdx.x == 0) slot = atomicAdd(counter, 1u);
    __syncthreads();
    return static_cast<int64_t>(slot);
}

// In-place Chen step on S (contiguous, global), x in LDS.  P0/P1 hold the
// Horner intermediates of one target level (width^(depth-1) each).
template <typename T, int W>
__device__ void chen_step(const DevBasis& B, T* S, const T* xs, const T* inv, T* P0, T* P1) {
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    for (int k = N; k >= 1; --k) {
        const T a0 = S[0];
        const T* prev = nullptr;
        for (int i = 1; i <= k; ++i) {
            const T c = inv[k - i + 1];
            T* dst = (i == k) ? S + B.start[k] : ((i & 1) ? P0 : P1);
            const T* Si = S + B.start[i];
            const uint32_t ni = B.pow[i];
            for (uint32_t w = threadIdx.x; w < ni; w += blockDim.x) {
                const uint32_t p = L::div(B, w);
                const uint32_t a = w - p * width;
                const T ap = (i == 1) ? a0 : prev[p];
                dst[w] = ap * (xs[a] * c) + Si[w];
            }
            __syncthreads();
            prev = dst;
        }
    }
}

// Reverse mode of one Chen step taken at state S (the state before the
// step, x in LDS): G, the cotangent of the new state, becomes in place the
// cotangent of S (degree 0 included), and part[thread * 16 + a] accumulates
// this thread's share of xbar[a] (cleared here).  Ast: Horner values of
// levels 0..depth-1; Q0/Q1: reverse ping-pong of width^(depth-1) each.
template <typename T, int W>
__device__ void chen_step_adjoint(const DevBasis& B, const T* S, T* G, const T* xs, const T* inv, T* Ast, T* Q0,
                                  T* Q1, T* part) {
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    for (uint32_t a = 0; a < width; ++a) part[threadIdx.x * kMaxBackwardWidth + a] = T(0);
    __syncthreads();
    for (int k = 1; k <= N; ++k) {
        // forward Horner values A_0..A_{k-1} of target k
        if (threadIdx.x == 0) Ast[0] = S[0];
        __syncthreads();
        for (int i = 1; i < k; ++i) {
            const T c = inv[k - i + 1];
            const uint32_t ni = B.pow[i];
            T* Ai = Ast + B.start[i];
            const T* Ap = Ast + B.start[i - 1];
            const T* Si = S + B.start[i];
            for (uint32_t w = threadIdx.x; w < ni; w += blockDim.x) {
                const uint32_t q = L::div(B, w);
                Ai[w] = Ap[q] * (xs[w - q * width] * c) + Si[w];
            }
            __syncthreads();
        }
        // reverse: children p*d + a of every prefix p, one thread each.  G_k
        // is read before any later target changes it (targets ascend and
        // target k only adds into levels below k).
        const T* src = G + B.start[k];
        for (int i = k; i >= 1; --i) {
            const T c = inv[k - i + 1];
            const uint32_t np = B.pow[i - 1];
            T* dst = (i & 1) ? Q0 : Q1;
            const T* Ap = Ast + B.start[i - 1];
            T* Gp = G + B.start[i - 1];
            for (uint32_t q = threadIdx.x; q < np; q += blockDim.x) {
                const T apc = Ap[q] * c;
                T acc = T(0);
                for (uint32_t a = 0; a < width; ++a) {
                    const T g = src[q * width + a];
                    part[threadIdx.x * kMaxBackwardWidth + a] += g * apc;
                    acc += g * xs[a];
                }
                acc *= c;
                dst[q] = acc;
                Gp[q] += acc;
            }
            __syncthreads();
            src = dst;
        }
    }
}

template <typename T, int W>
__global__ void __launch_bounds__(kSigBlock)
k_sig_fwd_generic(DevBasis B, int64_t n_paths, const int64_t* order, const T* incr,
                  const int64_t* offs, View<T> out, int ck_every, T* ckpts, const int64_t* ck_offs,
                  T* scratch, int64_t slot_elems, unsigned* counter) {
    __shared__ T xs[256];
    __shared__ T inv[kMaxDepth + 2];
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    const int64_t dim = B.start[N + 1];
    if (threadIdx.x <= static_cast<unsigned>(N + 1))
        inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
    T* P0 = scratch + static_cast<int64_t>(blockIdx.x) * slot_elems;
    T* P1 = P0 + (N >= 1 ? B.pow[N - 1] : 1);
    T* S = P1 + (N >= 1 ? B.pow[N - 1] : 1);
    for (;;) {
        const int64_t item = next_item(counter);
        if (item >= n_paths) break;
        const int64_t p = order ? order[item] : item;
        for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) S[i] = i == 0 ? T(1) : T(0);
        const int64_t t0 = offs[p], t1 = offs[p + 1];
        for (int64_t t = t0; t < t1; ++t) {
            __syncthreads();
            for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) xs[a] = incr[t * width + a];
            __syncthreads();
            chen_step<T, W>(B, S, xs, inv, P0, P1);
            if (ck_every > 0 && ((t - t0 + 1) % ck_every) == 0) {
                T* dst = ckpts + (ck_offs[p] + (t - t0 + 1) / ck_every - 1) * dim;
                for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) dst[i] = S[i];
            }
        }
        __syncthreads();
        T* o = out.obj(p);
        for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) o[i * out.cs] = S[i];
    }
}

template <typename T, int W>
__global__ void __launch_bounds__(kSigBlock)
k_sig_bwd_generic(DevBasis B, int64_t n_paths, const int64_t* order, const T* incr,
                  const int64_t* offs, View<const T> sigs, View<const T> cot, T* grad, int ck_every,
                  const T* ckpts, const int64_t* ck_offs, T* scratch, int64_t slot_elems,
                  unsigned* counter) {
    __shared__ T xs[kMaxBackwardWidth];
    __shared__ T xneg[kMaxBackwardWidth];
    __shared__ T inv[kMaxDepth + 2];
    __shared__ T part[kSigBlock * kMaxBackwardWidth];
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    const int64_t dim = B.start[N + 1];
    const uint32_t top = N >= 1 ? B.pow[N - 1] : 1;
    if (threadIdx.x <= static_cast<unsigned>(N + 1))
        inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
    T* S = scratch + static_cast<int64_t>(blockIdx.x) * slot_elems;
    T* G = S + dim;
    T* Ast = G + dim;                  // forward Horner values, levels 0..N-1
    T* Q0 = Ast + B.start[N];          // reverse ping-pong
    T* Q1 = Q0 + top;
    T* P0 = Q1 + top;                  // reconstruction scratch
    T* P1 = P0 + top;
    for (;;) {
        const int64_t item = next_item(counter);
        if (item >= n_paths) break;
        const int64_t p = order ? order[item] : item;
        const T* s_in = sigs.obj(p);
        const T* g_in = cot.obj(p);
        for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) {
            S[i] = s_in[i * sigs.cs];
            G[i] = g_in[i * cot.cs];
        }
        const int64_t t0 = offs[p], t1 = offs[p + 1];
        for (int64_t t = t1 - 1; t >= t0; --t) {
            __syncthreads();
            for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) {
                const T v = incr[t * width + a];
                xs[a] = v;
                xneg[a] = -v;
            }
            __syncthreads();
            // S <- S_{t-1}: exact at the start, reloaded at checkpoints,
            // otherwise reconstructed by S_t (x) exp(-x_t).
            const int64_t tl = t - t0;
            if (tl == 0) {
                for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) S[i] = i == 0 ? T(1) : T(0);
            } else if (ck_every > 0 && (tl % ck_every) == 0) {
                const T* src = ckpts + (ck_offs[p] + tl / ck_every - 1) * dim;
                for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) S[i] = src[i];
            } else {
                chen_step<T, W>(B, S, xneg, inv, P0, P1);
            }
            chen_step_adjoint<T, W>(B, S, G, xs, inv, Ast, Q0, Q1, part);
            // deterministic reduction of the per-thread partials
            for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) {
                T acc = T(0);
                for (unsigned j = 0; j < blockDim.x; ++j) acc += part[j * kMaxBackwardWidth + a];
                grad[t * width + a] = acc;
            }
        }
    }
}

// Per-operation adjoint of ft_fmexp with a degree-one exponent, any tile
// shape: one object per work-group iteration, S and G staged in the
// work-group's workspace slice (windows zero-extended).
template <typename T, int W>
__global__ void __launch_bounds__(kSigBlock)
k_fmexp_adj_generic(DevBasis B, int64_t n, View<const T> mult, View<const T> ex, View<const T> cot, View<T> gmult,
                    View<T> gx, T* scratch, int64_t slot_elems) {
    __shared__ T xs[kMaxBackwardWidth];
    __shared__ T inv[kMaxDepth + 2];
    __shared__ T part[kSigBlock * kMaxBackwardWidth];
    using L = Letters<W>;
    const uint32_t width = L::width(B);
    const int N = B.depth;
    const int64_t dim = B.start[N + 1];
    const uint32_t top = N >= 1 ? B.pow[N - 1] : 1;
    if (threadIdx.x <= static_cast<unsigned>(N + 1))
        inv[threadIdx.x] = threadIdx.x == 0 ? T(1) : T(1) / T(threadIdx.x);
    T* S = scratch + static_cast<int64_t>(blockIdx.x) * slot_elems;
    T* G = S + dim;
    T* Ast = G + dim;
    T* Q0 = Ast + B.start[N];
    T* Q1 = Q0 + top;
    const int64_t m0 = B.start[mult.mn], m1 = B.start[mult.mx + 1];
    const int64_t c0 = B.start[cot.mn], c1 = B.start[cot.mx + 1];
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        __syncthreads();
        const T* so = mult.obj(b);
        const T* go = cot.obj(b);
        for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) {
            S[i] = (m0 <= i && i < m1) ? so[i * mult.cs] : T(0);
            G[i] = (c0 <= i && i < c1) ? go[i * cot.cs] : T(0);
        }
        for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) xs[a] = ex.obj(b)[(1 + a) * ex.cs];
        __syncthreads();
        chen_step_adjoint<T, W>(B, S, G, xs, inv, Ast, Q0, Q1, part);
        T* xo = gx.obj(b);
        for (uint32_t a = threadIdx.x; a < width; a += blockDim.x) {
            T acc = T(0);
            for (unsigned j = 0; j < blockDim.x; ++j) acc += part[j * kMaxBackwardWidth + a];
            xo[(1 + a) * gx.cs] = acc;
        }
        if (threadIdx.x == 0 && gx.mn == 0) xo[0] = T(0);
        T* o = gmult.obj(b);
        for (int64_t i = threadIdx.x; i < dim; i += blockDim.x) o[i * gmult.cs] = G[i];
    }
}

#define RPP_SIG_DISPATCH(WIDTH, CALL)                                    \
    do {                                                                 \
        switch (WIDTH) {                                                 \
            case 2: { constexpr int W_ = 2; CALL; } break;               \
            case 3: { constexpr int W_ = 3; CALL; } break;               \
            case 4: { constexpr int W_ = 4; CALL; } break;               \
            case 5: { constexpr int W_ = 5; CALL; } break;               \
            case 6: { constexpr int W_ = 6; CALL; } break;               \
            case 8: { constexpr int W_ = 8; CALL; } break;               \
            default: { constexpr int W_ = 0; CALL; } break;              \
        }                                                                \
    } while (0)

}  // namespace

// Queue order: paths by decreasing length (longest first balances the tail).
int path_order(const int64_t* d_offsets, int64_t n, std::vector<int64_t>& lengths_host,
               int64_t** d_order, hipStream_t stream, int64_t* last_offset) {
    std::vector<int64_t> offs(static_cast<size_t>(n) + 1);
    if (int rc = check_hip(hipMemcpyAsync(offs.data(), d_offsets, offs.size() * sizeof(int64_t),
                                          hipMemcpyDeviceToHost, stream),
                           "copying path offsets"))
        return rc;
    if (int rc = check_hip(hipStreamSynchronize(stream), "synchronising for path offsets")) return rc;
    if (offs[0] < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "offsets[0] must be >= 0");
    if (last_offset) *last_offset = offs[static_cast<size_t>(n)];
    lengths_host.resize(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        lengths_host[i] = offs[i + 1] - offs[i];
        if (lengths_host[i] < 0) return fail(RPP_HIP_INVALID_ARGUMENT, "offsets must be non-decreasing");
    }
    std::vector<int64_t> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return lengths_host[a] > lengths_host[b]; });
    void* ws = nullptr;
    if (int rc = workspace(kSlotOrder, static_cast<size_t>(n) * sizeof(int64_t) + 256, &ws)) return rc;
    *d_order = static_cast<int64_t*>(ws);
    // The host vector dies at return: copy synchronously.
    return check_hip(hipMemcpy(*d_order, order.data(), order.size() * sizeof(int64_t), hipMemcpyHostToDevice),
                     "uploading the path order");
}

template <typename T>
int signature_forward_generic(const DevBasis& B, const rpp_hip_signature_desc& d, const int64_t* d_order,
                              hipStream_t stream) {
    const int N = B.depth;
    const int64_t top = N >= 1 ? B.pow[N - 1] : 1;
    const int64_t slot = 2 * top + B.start[N + 1];
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(d.num_paths, device_cu_count() * 4));
    void* ws = nullptr;
    if (int rc = workspace(kSlotState, static_cast<size_t>(slot) * grid * sizeof(T) + 256, &ws)) return rc;
    void* cnt = nullptr;
    if (int rc = workspace(kSlotCounter, 256, &cnt)) return rc;
    k_reset_counter<<<1, 64, 0, stream>>>(static_cast<unsigned*>(cnt));
    const View<T> out = make_view<T>(d.out);
    RPP_SIG_DISPATCH(B.width, (k_sig_fwd_generic<T, W_><<<grid, kSigBlock, 0, stream>>>(
                                  B, d.num_paths, d_order, static_cast<const T*>(d.increments), d.offsets, out,
                                  d.checkpoint_every, static_cast<T*>(d.checkpoints), d.ckpt_offsets,
                                  static_cast<T*>(ws), slot, static_cast<unsigned*>(cnt))));
    return check_hip(hipGetLastError(), "k_sig_fwd_generic launch");
}

int64_t backward_slot_elems(const DevBasis& B) {
    const int N = B.depth;
    const int64_t top = N >= 1 ? B.pow[N - 1] : 1;
    return 2 * B.start[N + 1] + B.start[N] + 4 * top;
}

template <typename T>
int signature_backward_generic(const DevBasis& B, const rpp_hip_signature_backward_desc& d,
                               const int64_t* d_order, hipStream_t stream) {
    if (B.width > kMaxBackwardWidth)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "signature backward supports width <= 16");
    const int64_t slot = backward_slot_elems(B);
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(d.num_paths, device_cu_count() * 4));
    void* ws = d.scratch;
    if (ws == nullptr) {
        if (int rc = workspace(kSlotState, static_cast<size_t>(slot) * grid * sizeof(T) + 256, &ws)) return rc;
    }
    void* cnt = nullptr;
    if (int rc = workspace(kSlotCounter, 256, &cnt)) return rc;
    k_reset_counter<<<1, 64, 0, stream>>>(static_cast<unsigned*>(cnt));
    const View<const T> sigs = make_cview<T>(d.signatures);
    const View<const T> cot = make_cview<T>(d.cotangent);
    RPP_SIG_DISPATCH(B.width, (k_sig_bwd_generic<T, W_><<<grid, kSigBlock, 0, stream>>>(
                                  B, d.num_paths, d_order, static_cast<const T*>(d.increments), d.offsets, sigs,
                                  cot, static_cast<T*>(d.grad_increments), d.checkpoint_every,
                                  static_cast<const T*>(d.checkpoints), d.ckpt_offsets, static_cast<T*>(ws), slot,
                                  static_cast<unsigned*>(cnt))));
    return check_hip(hipGetLastError(), "k_sig_bwd_generic launch");
}

template <typename T>
int fmexp_adjoint_generic(const DevBasis& B, const rpp_hip_ft_fmexp_adjoint_desc& d, hipStream_t stream) {
    if (B.width > kMaxBackwardWidth)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "ft_fmexp adjoint supports width <= 16 without a compiled tile");
    if (d.num_batches <= 0) return 0;
    const int64_t slot = backward_slot_elems(B);
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(d.num_batches, device_cu_count() * 4));
    void* ws = nullptr;
    if (int rc = workspace(kSlotState, static_cast<size_t>(slot) * grid * sizeof(T) + 256, &ws)) return rc;
    RPP_SIG_DISPATCH(B.width, (k_fmexp_adj_generic<T, W_><<<grid, kSigBlock, 0, stream>>>(
                                  B, d.num_batches, make_cview<T>(d.multiplier), make_cview<T>(d.exponent),
                                  make_cview<T>(d.cotangent), make_view<T>(d.grad_multiplier),
                                  make_view<T>(d.grad_exponent), static_cast<T*>(ws), slot)));
    return check_hip(hipGetLastError(), "k_fmexp_adj_generic launch");
}

template int fmexp_adjoint_generic<float>(const DevBasis&, const rpp_hip_ft_fmexp_adjoint_desc&, hipStream_t);
template int fmexp_adjoint_generic<double>(const DevBasis&, const rpp_hip_ft_fmexp_adjoint_desc&, hipStream_t);
template int signature_forward_generic<float>(const DevBasis&, const rpp_hip_signature_desc&, const int64_t*,
                                              hipStream_t);
template int signature_forward_generic<double>(const DevBasis&, const rpp_hip_signature_desc&, const int64_t*,
                                               hipStream_t);
template int signature_backward_generic<float>(const DevBasis&, const rpp_hip_signature_backward_desc&,
                                               const int64_t*, hipStream_t);
template int signature_backward_generic<double>(const DevBasis&, const rpp_hip_signature_backward_desc&,
                                                const int64_t*, hipStream_t);

}  // namespace rpp_hip
