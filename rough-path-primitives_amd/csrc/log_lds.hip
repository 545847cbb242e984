// ft_log with the object resident in LDS (operations/intermediate/ft_log.hpp:66-85).
//
// The reference's Horner scheme: out = 0; for d = N..1: out += c_d 1, then
// out <- out (x) Y on degrees [0, N-d+1], c_d = (-1)^(d+1) / d, Y = arg on
// degrees >= 1.  Here one work-group holds Y and out for one object in LDS
// (2 dim values) and each step computes every new coefficient it owns into
// registers from the old out and Y, then writes them back: two barriers per
// step instead of one per degree, and HBM sees one read of arg and one write
// of out per object.  The shape is fixed at compile time, so every degree
// offset and division by width^j is a constant; the split order (ascending j)
// is that of the product kernels.
#include "common.hpp"
#include "log_lds.hpp"

#include <algorithm>

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
    static constexpr int slots(int k, int nt) { return (words(k) + nt - 1) / nt; }
    static constexpr int reg_offset(int k, int nt) {
        int o = 0;
        for (int i = 1; i < k; ++i) o += slots(i, nt);
        return o;
    }
};

// 1024 threads unless one object's coefficients would then need more than 8
// registers per thread (the 128-register budget of a 1024-thread group
// would spill); 512 threads get 256 registers each.
template <typename T, int W, int N>
constexpr int log_threads() {
    return LogGeo<W, N>::dim * sizeof(T) > 8 * 1024 * 8 ? 512 : 1024;
}

template <typename T, int W, int N, int NT>
__global__ void __launch_bounds__(NT)
k_log_lds(View<T> out, View<const T> arg, int64_t n) {
    using G = LogGeo<W, N>;
    extern __shared__ __align__(16) unsigned char log_raw[];
    T* ys = reinterpret_cast<T*>(log_raw);
    T* os = ys + G::dim;
    const int t = threadIdx.x;
    const int ylo = max(arg.mn, 1), yhi = min(arg.mx, N);
    T reg[G::reg_offset(N + 1, NT) > 0 ? G::reg_offset(N + 1, NT) : 1];
    for (int64_t b = blockIdx.x; b < n; b += gridDim.x) {
        const T* a = arg.obj(b);
        __syncthreads();  // the previous object's readers are done
        for (int i = t; i < G::dim; i += NT) {
            int k = 0;
#pragma unroll
            for (int j = 1; j <= N; ++j)
                if (i >= G::start(j)) k = j;
            ys[i] = (k >= ylo && k <= yhi) ? a[i] : T(0);
            os[i] = T(0);
        }
        if (t == 0) os[0] = (N % 2 == 0 ? T(-1) : T(1)) / T(N);
        __syncthreads();
#pragma unroll 1
        for (int d = N; d >= 1; --d) {
            const int K = N - d + 1;
            // new out_k = sum_{j=1..k} P_{k-j} (x) Y_j for 1 <= k <= K, P = out
#pragma unroll
            for (int k = 1; k <= N; ++k) {
                if (k > K) continue;
#pragma unroll
                for (int s = 0; s < G::slots(k, NT); ++s) {
                    const int w = t + s * NT;
                    if (G::words(k) % NT == 0 || w < G::words(k)) {
                        T acc = T(0);
#pragma unroll
                        for (int j = 1; j <= k; ++j) {
                            if (j >= ylo && j <= yhi) {
                                const int p = w / G::words(j);
                                acc += os[G::start(k - j) + p] * ys[G::start(j) + (w - p * G::words(j))];
                            }
                        }
                        reg[G::reg_offset(k, NT) + s] = acc;
                    }
                    // keep the scheduler from hoisting every slot's LDS reads at
                    // once: the in-flight values would not fit in registers
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 1; k <= N; ++k) {
                if (k > K) continue;
#pragma unroll
                for (int s = 0; s < G::slots(k, NT); ++s) {
                    const int w = t + s * NT;
                    if (G::words(k) % NT == 0 || w < G::words(k)) os[G::start(k) + w] = reg[G::reg_offset(k, NT) + s];
                }
            }
            // the next step's identity term (degree 0 of the product is 0)
            if (t == 0) os[0] = d > 1 ? ((d - 1) % 2 == 0 ? T(-1) : T(1)) / T(d - 1) : T(0);
            __syncthreads();
        }
        T* o = out.obj(b);
        for (int i = t; i < G::dim; i += NT) o[i] = os[i];
    }
}

template <typename T, int W, int N>
int log_launch(int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg, hipStream_t stream) {
    constexpr size_t lds = 2 * static_cast<size_t>(LogGeo<W, N>::dim) * sizeof(T);
    constexpr int NT = log_threads<T, W, N>();
    static const bool raised = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_log_lds<T, W, N, NT>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   static_cast<int>(lds)) == hipSuccess;
    if (!raised && lds > 64 * 1024) return fail(RPP_HIP_OUT_OF_RESOURCES, "cannot raise the LDS limit");
    const int64_t per_cu = std::max<int64_t>(1, static_cast<int64_t>((160 * 1024) / lds));
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(n, device_cu_count() * per_cu)));
    k_log_lds<T, W, N, NT><<<grid, NT, lds, stream>>>(make_view<T>(out), make_cview<T>(arg), n);
    return check_hip(hipGetLastError(), "k_log_lds launch");
}

template <typename T>
using LogFn = int (*)(int64_t, const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&, hipStream_t);

template <typename T>
LogFn<T> log_kernel(int width, int depth) {
// Only shapes whose object fits in 64 KiB of LDS: with two or more work-groups
// per compute unit the barrier-separated steps of one object overlap another's;
// one work-group per unit (width 3, depth 8 in FP64: 157 KiB) measured slower
// than the per-degree kernel.
#define RPP_LOG(W, N) \
    if (width == W && depth == N && 2 * LogGeo<W, N>::dim * sizeof(T) <= 64 * 1024) return &log_launch<T, W, N>;
    RPP_LOG(2, 8) RPP_LOG(2, 10) RPP_LOG(3, 6) RPP_LOG(3, 7) RPP_LOG(3, 8) RPP_LOG(4, 5) RPP_LOG(4, 6)
    RPP_LOG(5, 4) RPP_LOG(5, 5) RPP_LOG(6, 4) RPP_LOG(8, 4)
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
