// tensor_to_lie / lie_to_tensor over the packed Lie matrices (lie.hpp).
//
// The conversions multiply every object of a batch by the same sparse matrix
// (T2L or L2T).  A work-group stages the argument segments of G objects and
// the matrix's value table in LDS, then streams the matrix once for all G
// objects: groups of 8 lanes take a row each, read its packed entries
// consecutively, gather from LDS and reduce in a fixed order.  Matrix traffic
// from L2 is therefore 4 bytes per entry per G objects.  A call may take a
// slice of the rows (the matrix's offsets from row0) whose entries name only
// the columns [col0, col0 + arg.size), the argument segment staged.
#include "common.hpp"
#include "lie.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int kLieBlock = 1024;
constexpr int kLieLanes = 8;
constexpr int kLieTable = 256;
constexpr size_t kLieLdsMax = 160 * 1024;

template <typename T, int G>
__global__ void __launch_bounds__(kLieBlock)
k_lie_spmv(T* out, int64_t obs, int64_t ocs, int64_t obeg, const T* arg, int64_t abs_, int64_t acs, int64_t abeg,
           int64_t cols, int64_t col0, const uint32_t* packed, const int32_t* offsets, const T* table,
           int table_size, int64_t rows, int64_t n) {
    extern __shared__ __align__(16) unsigned char lie_raw[];
    T* tab = reinterpret_cast<T*>(lie_raw);
    T* xs = tab + kLieTable;
    const int t = threadIdx.x;
    const int lane = t % kLieLanes;
    const int grp = t / kLieLanes;
    constexpr int groups = kLieBlock / kLieLanes;
    for (int i = t; i < table_size; i += kLieBlock) tab[i] = table[i];
    for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * G; b0 < n; b0 += static_cast<int64_t>(gridDim.x) * G) {
        const int gn = static_cast<int>(min<int64_t>(G, n - b0));
        __syncthreads();
        for (int g = 0; g < gn; ++g) {
            const T* a = arg + (b0 + g) * abs_;
            for (int64_t i = t; i < cols; i += kLieBlock) xs[g * cols + i] = a[(abeg + i) * acs];
        }
        __syncthreads();
        for (int64_t row = grp; row < rows; row += groups) {
            const int32_t e0 = offsets[row], e1 = offsets[row + 1];
            T acc[G];
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] = T(0);
            // four entries per lane in flight: the loads of one step are all
            // issued before the first gather waits on them
            constexpr int U = 4;
            for (int32_t e = e0 + lane; e < e1; e += U * kLieLanes) {
                uint32_t u[U];
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    const int32_t ej = e + j * kLieLanes;
                    u[j] = ej < e1 ? packed[ej] : 0u;
                }
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    if (e + j * kLieLanes < e1) {
                        const T v = tab[u[j] >> 24];
                        const uint32_t c = (u[j] & 0xffffffu) - static_cast<uint32_t>(col0);
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            if (g < gn) acc[g] += v * xs[g * cols + c];
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
#pragma unroll
                for (int off = kLieLanes / 2; off > 0; off >>= 1) acc[g] += __shfl_down(acc[g], off, kLieLanes);
            }
            if (lane == 0) {
#pragma unroll
                for (int g = 0; g < G; ++g)
                    if (g < gn) out[(b0 + g) * obs + (obeg + row) * ocs] = acc[g];
            }
        }
    }
}

template <typename T, int G>
int lie_launch(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
               int64_t row0, int64_t rows, int64_t col0, size_t lds, hipStream_t stream) {
    static const bool raised = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lie_spmv<T, G>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   static_cast<int>(kLieLdsMax)) == hipSuccess;
    if (!raised && lds > 64 * 1024) return fail(RPP_HIP_OUT_OF_RESOURCES, "cannot raise the LDS limit");
    const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(4, static_cast<int64_t>(kLieLdsMax / lds)));
    const int64_t groups = (n + G - 1) / G;
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min(groups, device_cu_count() * per_cu)));
    k_lie_spmv<T, G><<<grid, kLieBlock, lds, stream>>>(
        static_cast<T*>(out.data), out.batch_stride, out.coeff_stride, out.begin, static_cast<const T*>(arg.data),
        arg.batch_stride, arg.coeff_stride, arg.begin, arg.size, col0, m.packed, m.offsets32 + row0,
        static_cast<const T*>(m.table), m.table_size, rows, n);
    return check_hip(hipGetLastError(), "k_lie_spmv launch");
}

}  // namespace

template <typename T>
int launch_lie_spmv(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
                    int64_t row0, int64_t rows, int64_t col0, hipStream_t stream) {
    if (n <= 0 || rows <= 0) return 0;
    if (m.table_size == 0 || row0 < 0 || row0 + rows > m.rows || col0 < 0 || col0 + arg.size > m.cols)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "no packed Lie matrix for this row and column range");
    const size_t per_object = static_cast<size_t>(std::max<int64_t>(arg.size, 1)) * sizeof(T);
    const size_t fixed = kLieTable * sizeof(T);
    const int64_t fit = static_cast<int64_t>((kLieLdsMax - fixed) / per_object);
    if (fit < 1) return fail(RPP_HIP_NOT_IMPLEMENTED, "Lie argument does not fit in LDS");
    const int G = static_cast<int>(std::min<int64_t>({fit, 4, n}));
    const size_t lds = fixed + static_cast<size_t>(G) * per_object;
    switch (G) {
        case 1: return lie_launch<T, 1>(n, out, arg, m, row0, rows, col0, lds, stream);
        case 2: return lie_launch<T, 2>(n, out, arg, m, row0, rows, col0, lds, stream);
        case 3: return lie_launch<T, 3>(n, out, arg, m, row0, rows, col0, lds, stream);
        default: return lie_launch<T, 4>(n, out, arg, m, row0, rows, col0, lds, stream);
    }
}

template <typename T>
bool lie_spmv_fits(const DeviceCsr& m, const rpp_hip_vector_batch& arg) {
    const size_t per_object = static_cast<size_t>(std::max<int64_t>(arg.size, 1)) * sizeof(T);
    return m.table_size > 0 && kLieTable * sizeof(T) + per_object <= kLieLdsMax;
}

template int launch_lie_spmv<float>(int64_t, const rpp_hip_vector_batch&, const rpp_hip_vector_batch&,
                                    const DeviceCsr&, int64_t, int64_t, int64_t, hipStream_t);
template int launch_lie_spmv<double>(int64_t, const rpp_hip_vector_batch&, const rpp_hip_vector_batch&,
                                     const DeviceCsr&, int64_t, int64_t, int64_t, hipStream_t);
template bool lie_spmv_fits<float>(const DeviceCsr&, const rpp_hip_vector_batch&);
template bool lie_spmv_fits<double>(const DeviceCsr&, const rpp_hip_vector_batch&);

}  // namespace rpp_hip
