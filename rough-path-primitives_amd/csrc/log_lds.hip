// ft_log with the object's lower degrees resident in LDS
// (operations/intermediate/ft_log.hpp:66-85).
//
// The reference's Horner scheme: P = 0; for d = N..1: P += c_d 1, then
// P <- P (x) Y on degrees [0, N-d+1], c_d = (-1)^(d+1) / d, Y = arg on
// degrees >= 1; out = P.  Only the last step (d = 1) reaches degree N, and
// there the top degrees of both factors enter element-wise: out_N[w] =
// sum_{j<N} P_{N-j}[w / d^j] Y_j[w mod d^j] + P_0 Y_N[w], and nothing reads
// out_N afterwards.  So one work-group keeps only degrees 0..N-1 of Y and of
// P in LDS (about 1/width of the object each), reads Y_N from HBM at the last
// step and writes every degree of that step to HBM directly.  P is kept twice:
// each earlier step reads one copy and writes its new coefficients into the
// other, so a step ends with one barrier and holds nothing in registers.  HBM
// sees one read of arg and one write of out per object.  A work-group loads
// the lower degrees of its next object into registers as soon as the current
// one is staged, so that read overlaps the Horner steps, and the last step
// reads each top-degree argument word before the LDS reads of its sum.  The
// shape is fixed at compile time, so every degree offset and division by
// width^j is a constant.  Each sum adds its terms in the order of the
// reference's CPU product (cpu_term_no_rhs0: ascending rhs degree below the
// first right-tiled degree, the right-tiled order at and above it), every
// product and sum rounded on its own (contraction off), so the logarithm
// agrees with the reference's cpu::single_thread ft_log bit for bit.
//
// Width 3, depth 8 over 8256 objects on MI355X, microseconds per call: FP64
// 558 with P updated in place (two barriers per step, new values staged in
// registers, one work-group per compute unit) -> 339 double-buffered (75
// registers, held to 64 so that two work-groups share a compute unit); FP32
// 630 -> 498.  Every one of the 16 compiled shapes gains, 1-50 %.  In FP32 a
// thread sums blocks of width consecutive words at most shapes, sharing each
// P read among them (log_block_instance).
#include "common.hpp"
#include "log_lds.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rpp_hip {

namespace {

constexpr int lpow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

template <int W, int N>
struct LogGeo {
    static constexpr int words(int k) { return lpow(W, k); }
    static constexpr int start(int k) {
        int s = 0;
        for (int i = 0; i < k; ++i) s += lpow(W, i);
        return s;
    }
    static constexpr int dim = start(N + 1);
    static constexpr int low = start(N);  // degrees 0..N-1, the LDS-resident part
    static constexpr int slots(int k, int nt) { return (words(k) + nt - 1) / nt; }
};

// LDS of one work-group: Y and two copies of P on degrees 0..N-1.
template <typename T, int W, int N>
constexpr size_t log_lds_bytes() {
    return 3 * static_cast<size_t>(LogGeo<W, N>::low) * sizeof(T);
}

constexpr size_t kLogLdsMax = 160 * 1024;

// Work-groups of a compute unit by LDS, and the waves per SIMD that many
// work-groups of nt threads make (at most 8).
template <typename T, int W, int N>
constexpr int log_groups_per_cu() {
    return static_cast<int>(kLogLdsMax / log_lds_bytes<T, W, N>());
}
template <typename T, int W, int N>
constexpr int log_waves(int nt) {
    const int w = log_groups_per_cu<T, W, N>() * nt / 256;
    return w < 8 ? w : 8;
}

// 512 threads where LDS still allows 8 waves per SIMD with them, else 1024;
// the kernel is compiled for the waves that LDS allows, so its registers do
// not cap the occupancy.  Measured over the 16 compiled shapes in both types
// (three candidates each: 512 or 1024 threads double-buffered, and the
// in-place form): this rule takes the fastest in 31 of 32 cases, and is 3 %
// behind at width 6 depth 5 FP64 (261 against 253 us with 1024 threads).
template <typename T, int W, int N>
constexpr int log_threads() {
    return log_waves<T, W, N>(512) >= 8 ? 512 : 1024;
}

// sum_u P_{K-j}[w / width^j] * Y_j[w mod width^j] over the terms u = 0 ..
// K-1 of output degree K in the reference's product order (j its u-th rhs
// degree), Y restricted to [ylo, yhi]; every index is a compile-time
// constant.  TOP: the term j = K = N reads Y_N from y (from HBM) with P_0 = p0.
template <typename T, int W, int N, int K, bool TOP>
__device__ __forceinline__ T log_level_sum(const T* ps, const T* ys, int w, int ylo, int yhi, T p0, T y, bool top) {
    using G = LogGeo<W, N>;
    constexpr int kF = cpu_first_tiled_degree(W), kR = cpu_right_letters(W);
    T acc = T(0);
    static_for<K>([&](auto uc) {
        constexpr int j = cpu_term_no_rhs0_fr(kF, kR, W == 1, K, decltype(uc)::value);
        if constexpr (TOP && j == K) {
            if (top) acc += p0 * y;
        } else {
            constexpr int wj = G::words(j), sk = G::start(K - j), sj = G::start(j);
            if (j >= ylo && j <= yhi) {
                const int p = w / wj;
                acc += ps[sk + p] * ys[sj + (w - p * wj)];
            }
        }
    });
    return acc;
}

// The sums of log_level_sum for the W consecutive words W m .. W m + W - 1 of
// degree K, each in the same term order: for j >= 1 the W words share
// P_{K-j}[m / width^(j-1)] and read Y_j[width (m mod width^(j-1)) + b], so one
// P read serves W words.  TOP: the term j = K = N adds p0 * y[b].
template <typename T, int W, int N, int K, bool TOP>
__device__ __forceinline__ void log_block_sum(T (&acc)[W], const T* ps, const T* ys, int m, int ylo, int yhi, T p0,
                                              const T (&y)[W], bool top) {
    using G = LogGeo<W, N>;
    constexpr int kF = cpu_first_tiled_degree(W), kR = cpu_right_letters(W);
#pragma unroll
    for (int b = 0; b < W; ++b) acc[b] = T(0);
    static_for<K>([&](auto uc) {
        constexpr int j = cpu_term_no_rhs0_fr(kF, kR, W == 1, K, decltype(uc)::value);
        if constexpr (TOP && j == K) {
            if (top) {
#pragma unroll
                for (int b = 0; b < W; ++b) acc[b] += p0 * y[b];
            }
        } else {
            constexpr int wj1 = G::words(j - 1), sk = G::start(K - j), sj = G::start(j);
            if (j >= ylo && j <= yhi) {
                const int p = m / wj1;
                const T pv = ps[sk + p];
                const T* yy = ys + sj + W * (m - p * wj1);
#pragma unroll
                for (int b = 0; b < W; ++b) acc[b] += pv * yy[b];
            }
        }
    });
}

// FULL: the argument window holds degrees 1..N, so its tests fold away.
// BLK: each thread sums blocks of W consecutive words (log_block_sum).
template <typename T, int W, int N, int NT, int WPE, bool FULL, bool BLK = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
k_log_lds(View<T> out, View<const T> arg, int64_t n) {
    using G = LogGeo<W, N>;
    extern __shared__ __align__(16) unsigned char log_raw[];
    T* ys = reinterpret_cast<T*>(log_raw);
    T* pa = ys + G::low;
    T* pb = pa + G::low;
    const int t = threadIdx.x;
    const int ylo = FULL ? 1 : max(arg.mn, 1), yhi = FULL ? N : min(arg.mx, N);
    const bool top = N >= ylo && N <= yhi;
    // degrees 0..N-1 of the object to stage next: loaded while the previous
    // object computes, so its HBM latency is hidden behind the Horner steps
    constexpr int SL = (G::low + NT - 1) / NT;
    T yv[SL];
    auto load_low = [&](int64_t b) {
        const T* a = arg.obj(b);
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int i = t + s * NT;
            int k = 0;
#pragma unroll
            for (int j = 1; j < N; ++j)
                if (i >= G::start(j)) k = j;
            yv[s] = (i < G::low && k >= ylo && k <= yhi) ? a[i] : T(0);
        }
    };
    int64_t b = blockIdx.x;
    if (b < n) load_low(b);
    for (; b < n; b += gridDim.x) {
        const T* a = arg.obj(b);
        __syncthreads();  // the previous object's readers are done
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int i = t + s * NT;
            if (i < G::low) ys[i] = yv[s];
        }
        // the first step reads only P_0 = c_N
        if (t == 0) pa[0] = (N % 2 == 0 ? T(-1) : T(1)) / T(N);
        __syncthreads();
        if (b + gridDim.x < n) load_low(b + gridDim.x);
        // steps d = N..2 reach degree K = N - d + 1 <= N - 1; each reads pc
        // on degrees 0..K-1 and writes degrees 1..K of pn, and the next step
        // reads pn on degrees 0..K only
        T* pc = pa;
        T* pn = pb;
#pragma unroll 1
        for (int d = N; d >= 2; --d) {
            const int K = N - d + 1;
            // new P_k = sum_{j=1..k} P_{k-j} (x) Y_j for 1 <= k <= K
            static_for<N>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if constexpr (k >= 1 && BLK) {
                    if (k <= K) {
                        constexpr int nb = G::words(k - 1);
#pragma unroll
                        for (int s = 0; s < (nb + NT - 1) / NT; ++s) {
                            const int m = t + s * NT;
                            if (nb % NT == 0 || m < nb) {
                                T acc[W];
                                const T yz[W] = {};
                                log_block_sum<T, W, N, k, false>(acc, pc, ys, m, ylo, yhi, T(0), yz, false);
#pragma unroll
                                for (int b = 0; b < W; ++b) pn[G::start(k) + W * m + b] = acc[b];
                            }
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                } else if constexpr (k >= 1) {
                    if (k <= K) {
#pragma unroll
                        for (int s = 0; s < G::slots(k, NT); ++s) {
                            const int w = t + s * NT;
                            if (G::words(k) % NT == 0 || w < G::words(k))
                                pn[G::start(k) + w] = log_level_sum<T, W, N, k, false>(pc, ys, w, ylo, yhi, T(0), T(0), false);
                            // keep the scheduler from hoisting every slot's LDS reads at
                            // once: the in-flight values would not fit in registers
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                }
            });
            // the next step's identity term (degree 0 of the product is 0)
            if (t == 0) pn[0] = ((d - 1) % 2 == 0 ? T(-1) : T(1)) / T(d - 1);
            __syncthreads();
            T* tmp = pc;
            pc = pn;
            pn = tmp;
        }
        // the last step (d = 1, P_0 = c_1 = 1) straight to HBM; degree 0 of
        // the product is 0.  Ordinary stores: non-temporal ones measured 1 %
        // faster at width 3 depth 8 over 16384 objects FP64 (0.573 -> 0.566
        // ms) but wrote 1.05 x the output bytes (partial lines at object
        // edges go to HBM unmerged)
        T* o = out.obj(b);
        if (t == 0) o[0] = T(0);
        static_for<N>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (k >= 1 && BLK) {
                for (int m = t; m < G::words(k - 1); m += NT) {
                    T acc[W];
                    const T yz[W] = {};
                    log_block_sum<T, W, N, k, false>(acc, pc, ys, m, ylo, yhi, T(0), yz, false);
#pragma unroll
                    for (int b = 0; b < W; ++b) o[G::start(k) + W * m + b] = acc[b];
                }
            } else if constexpr (k >= 1) {
                for (int w = t; w < G::words(k); w += NT)
                    o[G::start(k) + w] = log_level_sum<T, W, N, k, false>(pc, ys, w, ylo, yhi, T(0), T(0), false);
            }
        });
        const T p0 = pc[0];
        if constexpr (BLK) {
            for (int m = t; m < G::words(N - 1); m += NT) {
                // issued before the LDS reads of the sums, whose latency they overlap
                T y[W];
#pragma unroll
                for (int b = 0; b < W; ++b) y[b] = top ? a[G::start(N) + W * m + b] : T(0);
                T acc[W];
                log_block_sum<T, W, N, N, true>(acc, pc, ys, m, ylo, yhi, p0, y, top);
#pragma unroll
                for (int b = 0; b < W; ++b) o[G::start(N) + W * m + b] = acc[b];
            }
        } else {
            for (int w = t; w < G::words(N); w += NT) {
                // issued before the LDS reads of the sum, whose latency it overlaps
                const T y = top ? a[G::start(N) + w] : T(0);
                o[G::start(N) + w] = log_level_sum<T, W, N, N, true>(pc, ys, w, ylo, yhi, p0, y, top);
            }
        }
    }
}

template <typename T, int W, int N, int NT, int WPE, bool FULL, bool BLK = false>
int log_launch_with(int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg,
                    hipStream_t stream) {
    constexpr size_t lds = log_lds_bytes<T, W, N>();
    static_assert(lds <= kLogLdsMax, "the LDS of one compute unit");
    const auto kernel = &k_log_lds<T, W, N, NT, WPE, FULL, BLK>;
    static const bool raised = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   static_cast<int>(lds)) == hipSuccess;
    if (!raised && lds > 64 * 1024) return fail(RPP_HIP_OUT_OF_RESOURCES, "cannot raise the LDS limit");
    // objects are dealt to work-groups in turn, so the grid is what the
    // compute units hold at once (threads, registers and this LDS size)
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, NT, lds) != hipSuccess) resident = 1;
    const int64_t per_cu = std::max(1, resident);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(n, device_cu_count() * per_cu)));
    kernel<<<grid, NT, lds, stream>>>(make_view<T>(out), make_cview<T>(arg), n);
    return check_hip(hipGetLastError(), "k_log_lds launch");
}

// A full argument window (degrees 1..N) runs an instance whose window tests
// fold away.  Measured over the 16 compiled shapes in both types (microseconds
// per call, windowed / full): faster in 29 of 32 cases, e.g. width 3 depth 8
// FP64 323 / 273 and FP32 463 / 412, width 5 depth 6 FP64 330 / 261, width 2
// depth 12 FP64 858 / 481; slower where the full instance spills, FP64 width 2
// depth 10 (531 / 718) and width 4 depth 6 (236 / 251), which keep the
// windowed instance.
template <typename T, int W, int N>
constexpr bool log_full_instance() {
    return !(sizeof(T) == 8 && ((W == 2 && N == 10) || (W == 4 && N == 6)));
}

// Blocks of W words per thread (log_block_sum): one P read serves W words.
// Measured against one word per thread over the 16 compiled shapes (full
// window, microseconds per call): FP32 faster at 11 of them, e.g. width 3
// depth 8 427 -> 268, width 5 depth 6 316 -> 262, width 4 depth 6 271 -> 232,
// width 6 depth 4 366 -> 317; slower at widths 2, width 3 depth 6 and width 4
// depth 5 (e.g. width 2 depth 8 617 -> 731), which keep one word per thread;
// FP64 slower at every shape (width 3 depth 8 269 -> 287: its W sums and
// top-degree words in flight exceed the 64-register budget).
template <typename T, int W, int N>
constexpr bool log_block_instance() {
    return sizeof(T) == 4 && W >= 3 && !(W == 3 && N == 6) && !(W == 4 && N == 5);
}

template <typename T, int W, int N>
int log_launch(int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg, hipStream_t stream) {
    constexpr int NT = log_threads<T, W, N>();
    constexpr int WP = log_waves<T, W, N>(NT);
    constexpr bool BLK = log_block_instance<T, W, N>();
    if (log_full_instance<T, W, N>() && arg.min_degree <= 1 && arg.max_degree >= N)
        return log_launch_with<T, W, N, NT, WP, true, BLK>(n, out, arg, stream);
    return log_launch_with<T, W, N, NT, WP, false, BLK>(n, out, arg, stream);
}

template <typename T>
using LogFn = int (*)(int64_t, const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&, hipStream_t);

// The launcher of a shape whose LDS fits one compute unit, else nullptr.
template <typename T, int W, int N>
constexpr LogFn<T> log_fn() {
    if constexpr (log_lds_bytes<T, W, N>() <= kLogLdsMax) {
        return &log_launch<T, W, N>;
    } else {
        return nullptr;
    }
}

template <typename T>
LogFn<T> log_kernel(int width, int depth) {
#define RPP_LOG(W, N) \
    if (width == W && depth == N) return log_fn<T, W, N>();
    RPP_LOG(2, 8) RPP_LOG(2, 10) RPP_LOG(2, 12) RPP_LOG(3, 6) RPP_LOG(3, 7) RPP_LOG(3, 8) RPP_LOG(4, 5)
    RPP_LOG(4, 6) RPP_LOG(4, 7) RPP_LOG(5, 4) RPP_LOG(5, 5) RPP_LOG(5, 6) RPP_LOG(6, 4) RPP_LOG(6, 5)
    RPP_LOG(8, 4) RPP_LOG(8, 5)
#undef RPP_LOG
    return nullptr;
}

}  // namespace

template <typename T>
bool log_lds_available(int width, int depth) {
    return log_kernel<T>(width, depth) != nullptr;
}

template <typename T>
int launch_log_lds(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg,
                   hipStream_t stream) {
    const LogFn<T> fn = log_kernel<T>(B.width, B.depth);
    if (!fn) return fail(RPP_HIP_NOT_IMPLEMENTED, "no compiled LDS-resident ft_log for this shape");
    if (n <= 0) return 0;
    return fn(n, out, arg, stream);
}

template bool log_lds_available<float>(int, int);
template bool log_lds_available<double>(int, int);
template int launch_log_lds<float>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                   const rpp_hip_tensor_batch&, hipStream_t);
template int launch_log_lds<double>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                    const rpp_hip_tensor_batch&, hipStream_t);

}  // namespace rpp_hip
