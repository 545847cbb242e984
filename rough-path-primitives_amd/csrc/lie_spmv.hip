// tensor_to_lie / lie_to_tensor over the packed Lie matrices (lie.hpp).
//
// The conversions multiply every object of a batch by the same sparse matrix
// (T2L or L2T, block diagonal by degree).  A work-group stages the argument
// segments of OB objects in LDS, laid out [column][object], and each thread
// computes one (row, object) output as a sequential sum over the row's
// entries in ascending column order, acc = acc + v * x with contraction off.
// That is the reference's sum term by term, for a CSR matrix
// (cpu/.../linalg/sparse_matrix_vector.hpp:60-71) and for the CSC scatter of
// the builders' matrices (:146-160, which adds into each row in column
// order), so results agree bit for bit.
//
// Thread t serves object t % OB of row slot t / OB: the OB threads of a slot
// read the same entries (one address), and their argument reads are OB
// consecutive words.  Each row's entries are padded to a multiple of four
// (DeviceCsr::packed4) with entries that add (-0) * 0 = -0, which leaves
// every sum unchanged (x + -0 = x for every x, signed zeros, infinities and
// NaN included), so a row is read in 16-byte groups with no per-entry test.  Slots take the rows in the matrix's length order
// (longest first within each degree, DeviceCsr::rowdesc), so the rows of one
// wavefront have similar entry counts and the longest rows start first.  The
// sums are chains of dependent steps, so latency is hidden by many of them:
// 1024-thread work-groups holding at most 80 KiB of LDS where they can, two
// to a compute unit (one stages its objects while the other sums), and
// sixteen entries in flight per thread while sixteen are summed.
#include "common.hpp"
#include "lie.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rpp_hip {

namespace {

constexpr int kLieThreads = 1024;
constexpr size_t kLieLdsMax = 160 * 1024;
constexpr size_t kLieLdsShared = 80 * 1024;

template <typename T>
constexpr size_t lie_table_bytes(int table_size) {
    return (static_cast<size_t>(table_size) * sizeof(T) + 15) / 16 * 16;
}

// One packed entry: acc + table[w >> 24] * x[column].  The column is taken
// relative to the staged segment and clamped to its extra zero slot, where
// the padding entries (column 0xffffff) land.
// A thread serves J objects og * J .. og * J + J - 1 of its row: the entry is
// decoded and its table value read once for all of them.
template <typename T, int OB, int J>
__device__ __forceinline__ void lie_term(T (&acc)[J], uint32_t w, const T* tab, const T* xs, uint32_t col0,
                                         uint32_t cols, uint32_t og) {
    const uint32_t rel = min((w & 0xffffffu) - col0, cols);
    const T v = tab[w >> 24];
    const T* x = xs + rel * OB + og * J;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const T p = v * x[j];
        acc[j] = acc[j] + p;
    }
}

// Two work-groups of kLieThreads per compute unit: 8 waves per SIMD.
// A thread sums R rows at a time (their loads overlap) and keeps D groups of
// each in flight while D are summed.  Each thread serves J of the OB objects
// of its row, so OB / J threads share a row slot.  When the rows are enough
// to occupy the threads one to a slot (half a work-group or more), J = OB:
// the entry decoding, table read and LDS instruction of a term then serve
// every object.  (R, D) = (1, 4) for matrices with long rows, whose longest
// rows are a work-group's critical path, and (1, 2) for short rows
// (lie_to_tensor at width 3 depth 8 over 16384 objects: FP64 0.70 -> 0.46 ms,
// FP32 0.55 -> 0.26 ms against one object per thread).  With fewer rows, J =
// 1 and short-row matrices take R = 3 or 4 rows per thread, so that a row's
// two dependent loads (descriptor, entries) overlap those of the others.
// WPE: the waves per SIMD the register budget is set for (8 where two
// work-groups share a compute unit, 4 where one takes the whole LDS).
// UNIT: both coefficient strides are 1 (object-major batches), so no
// address needs a stride product.  PF: long-row matrices load their entry
// groups with no test (below).
template <typename T, int OB, int R, int D, int J, int WPE, bool UNIT, bool PF>
__global__ void __launch_bounds__(kLieThreads) __attribute__((amdgpu_waves_per_eu(WPE)))
k_lie_rows(T* out, int64_t obs, int64_t ocs_, int64_t obeg, const T* arg, int64_t abs_, int64_t acs_, int64_t abeg,
           uint32_t cols, uint32_t col0, const uint4* packed4, const int4* rowdesc, const T* table, int table_size,
           int64_t row0, int rows, int64_t n) {
    const int64_t ocs = UNIT ? 1 : ocs_, acs = UNIT ? 1 : acs_;
    extern __shared__ __align__(16) unsigned char lie_raw[];
    T* tab = reinterpret_cast<T*>(lie_raw);
    T* xs = reinterpret_cast<T*>(lie_raw + lie_table_bytes<T>(table_size));
    static_assert(OB % J == 0, "a row slot serves whole objects");
    constexpr int slack_last = static_cast<int>(kLiePackSlack) - 1;
    constexpr int kStage = kLieThreads / OB;        // staging steps per pass
    constexpr int kSlots = kLieThreads / (OB / J);  // row slots
    const uint32_t t = threadIdx.x;
    const uint32_t o = t % OB;  // staged object
    const uint32_t og = t % (OB / J);
    const int slot = static_cast<int>(t / (OB / J));
    for (int i = static_cast<int>(t); i < table_size; i += kLieThreads) tab[i] = table[i];
    for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * OB; b0 < n; b0 += static_cast<int64_t>(gridDim.x) * OB) {
        const uint32_t gn = static_cast<uint32_t>(min<int64_t>(OB, n - b0));
        __syncthreads();  // the previous group's sums have read xs
        // staged word c * OB + o (thread t stages object o = t % OB, which
        // 1024 / OB steps keep fixed): consecutive threads store consecutive
        // LDS words, and U loads are in flight before the first store (with
        // two work-groups per compute unit, 64 KiB in flight per CU)
        constexpr int U = 4;
        const T* src = arg + (b0 + static_cast<int64_t>(o)) * abs_ + abeg * acs;
        const bool live = o < gn;
        for (uint32_t c0 = t / OB; c0 <= cols; c0 += U * kStage) {
            T v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const uint32_t c = c0 + j * kStage;
                v[j] = (live && c < cols) ? src[static_cast<int64_t>(c) * acs] : T(0);
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const uint32_t c = c0 + j * kStage;
                if (c <= cols) xs[c * OB + o] = v[j];
            }
        }
        __syncthreads();
        for (int s0 = slot; s0 < rows; s0 += R * kSlots) {
            // {row, first group, end group} of each row: one load, no chain
            // of dependent index reads; each group holds four entries
            int4 d[R];
            int len[R];
            T acc[R][J];
            int most = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int s = s0 + r * kSlots;
                d[r] = s < rows ? rowdesc[row0 + s] : make_int4(0, 0, 0, 0);
                len[r] = d[r].z - d[r].y;
                most = max(most, len[r]);
#pragma unroll
                for (int j = 0; j < J; ++j) acc[r][j] = T(0);
            }
            // PF: every group load is unconditional: groups past a row's end
            // (the next rows', or the slack groups after the last row) are
            // loaded and never summed; an index is held below the row's end
            // group plus the slack, so it stays in the array.  With the loads
            // under a branch the compiler waits for the whole prefetch before
            // the current groups' sums (tensor_to_lie at width 3 depth 8 over
            // 16384 objects FP64: 1.12 -> 1.02 ms without the branch).  Short
            // rows keep the tests: most of their prefetches would pass the
            // row's end (lie_to_tensor 0.467 -> 0.475 ms without them).
            auto group = [&](int r, int i) {
                if constexpr (PF) {
                    return packed4[min(d[r].y + i, d[r].z + slack_last)];
                } else {
                    return i < len[r] ? packed4[d[r].y + i] : make_uint4(0, 0, 0, 0);
                }
            };
            uint4 u[R][D];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < D; ++j) u[r][j] = group(r, j);
            for (int k = 0; k < most; k += D) {
                uint4 nx[R][D];
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int j = 0; j < D; ++j) nx[r][j] = group(r, k + D + j);
#pragma unroll
                for (int r = 0; r < R; ++r) {
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        if (k + j < len[r]) {
                            lie_term<T, OB, J>(acc[r], u[r][j].x, tab, xs, col0, cols, og);
                            lie_term<T, OB, J>(acc[r], u[r][j].y, tab, xs, col0, cols, og);
                            lie_term<T, OB, J>(acc[r], u[r][j].z, tab, xs, col0, cols, og);
                            lie_term<T, OB, J>(acc[r], u[r][j].w, tab, xs, col0, cols, og);
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int j = 0; j < D; ++j) u[r][j] = nx[r][j];
            }
            // ordinary stores (non-temporal ones measured no faster for
            // tensor_to_lie and 4 % slower for lie_to_tensor at width 3 depth 8)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const uint32_t ob = og * J + j;
                    if (s0 + r * kSlots < rows && ob < gn) out[(b0 + ob) * obs + (obeg + (d[r].x - row0)) * ocs] = acc[r][j];
                }
        }
    }
}

template <typename T>
size_t lie_lds_bytes(const DeviceCsr& m, int64_t cols, int ob) {
    return lie_table_bytes<T>(m.table_size) + static_cast<size_t>(cols + 1) * ob * sizeof(T);
}

// Objects per work-group: the most (up to 8) whose staging fits in 80 KiB,
// so that two work-groups share a compute unit; else two objects, or one,
// in the whole LDS (one work-group per compute unit: each table read and
// each step of a long row then serves two objects, which measured faster
// than two one-object work-groups); 0 if not even one fits.
template <typename T>
int lie_objects_per_group(const DeviceCsr& m, int64_t cols) {
    for (int ob : {8, 4, 2})
        if (lie_lds_bytes<T>(m, cols, ob) <= kLieLdsShared) return ob;
    for (int ob : {2, 1})
        if (lie_lds_bytes<T>(m, cols, ob) <= kLieLdsMax) return ob;
    return 0;
}

// Long-row matrices with one thread per row for every object: the largest
// group (up to 8 objects) the whole LDS holds.  Larger groups share each
// entry's decoding among more objects, which measured faster than two
// work-groups of half the size per compute unit (tensor_to_lie at width 3
// depth 8 over 16384 objects, FP32: 4 objects 0.64 ms, two groups of 2 0.74).
template <typename T>
int lie_long_objects(const DeviceCsr& m, int64_t cols) {
    for (int ob : {8, 4, 2, 1})
        if (lie_lds_bytes<T>(m, cols, ob) <= kLieLdsMax) return ob;
    return 0;
}

template <typename T, int OB, int R, int D, int J, int WPE, bool UNIT, bool PF>
int lie_launch_with(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
                    int64_t row0, int64_t rows, int64_t col0, hipStream_t stream) {
    const auto kernel = &k_lie_rows<T, OB, R, D, J, WPE, UNIT, PF>;
    static const bool raised = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   static_cast<int>(kLieLdsMax)) == hipSuccess;
    const size_t lds = lie_lds_bytes<T>(m, arg.size, OB);
    if (!raised && lds > 64 * 1024) return fail(RPP_HIP_OUT_OF_RESOURCES, "cannot raise the LDS limit");
    // object groups are dealt to work-groups in turn, so the grid is what the
    // compute units hold at once (threads, registers and this LDS size)
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, kLieThreads, lds) != hipSuccess) resident = 1;
    const int64_t per_cu = std::max(1, resident);
    const int64_t groups = (n + OB - 1) / OB;
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min(groups, device_cu_count() * per_cu)));
    kernel<<<grid, kLieThreads, lds, stream>>>(
        static_cast<T*>(out.data), out.batch_stride, out.coeff_stride, out.begin, static_cast<const T*>(arg.data),
        arg.batch_stride, arg.coeff_stride, arg.begin, static_cast<uint32_t>(arg.size), static_cast<uint32_t>(col0),
        m.packed4, m.rowdesc, static_cast<const T*>(m.table), m.table_size, row0, static_cast<int>(rows), n);
    return check_hip(hipGetLastError(), "k_lie_rows launch");
}

template <typename T, int OB, int R, int D, int J = 1, int WPE = 8, bool PF = false>
int lie_launch(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
               int64_t row0, int64_t rows, int64_t col0, hipStream_t stream) {
    if (out.coeff_stride == 1 && arg.coeff_stride == 1)
        return lie_launch_with<T, OB, R, D, J, WPE, true, PF>(n, out, arg, m, row0, rows, col0, stream);
    return lie_launch_with<T, OB, R, D, J, WPE, false, PF>(n, out, arg, m, row0, rows, col0, stream);
}

}  // namespace

template <typename T>
int launch_lie_spmv(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
                    int64_t row0, int64_t rows, int64_t col0, hipStream_t stream) {
    if (n <= 0 || rows <= 0) return 0;
    if (m.table_size == 0 || !m.rowdesc || !m.packed4 || row0 < 0 || row0 + rows > m.rows || col0 < 0 || col0 + arg.size > m.cols)
        return fail(RPP_HIP_NOT_IMPLEMENTED, "no packed Lie matrix for this row and column range");
    const bool long_rows = m.long_rows != 0;
    // one thread per row for all objects of a group when the rows occupy at
    // least half of a work-group's threads that way
    const bool joint = rows >= kLieThreads / 2;
    // otherwise, rows per thread of short-row matrices: as many as fit in 64
    // registers
    constexpr int RS = sizeof(T) == 8 ? 3 : 4;
    if (long_rows && joint) {
        // one work-group per compute unit where the group takes more than
        // half the LDS: registers for 4 waves per SIMD, four entry groups in
        // flight while four are summed (width 3 depth 8, 16384 objects,
        // against six and eight: FP64 1.00 / 1.03 / 1.10 ms, FP32 0.657 /
        // 0.663 / 0.680 ms)
        const int ob = lie_long_objects<T>(m, arg.size);
        const bool whole = lie_lds_bytes<T>(m, arg.size, ob) > kLieLdsShared;
        switch (ob * (whole ? -1 : 1)) {
            // FP64 with 8 objects: two groups in flight fit in 64 registers
            case 8: return lie_launch<T, 8, 1, sizeof(T) == 8 ? 2 : 4, 8, 8, true>(n, out, arg, m, row0, rows, col0, stream);
            case 4: return lie_launch<T, 4, 1, 4, 4, 8, true>(n, out, arg, m, row0, rows, col0, stream);
            case 2: return lie_launch<T, 2, 1, 4, 2, 8, true>(n, out, arg, m, row0, rows, col0, stream);
            case 1: return lie_launch<T, 1, 1, 4, 1, 8, true>(n, out, arg, m, row0, rows, col0, stream);
            case -8: return lie_launch<T, 8, 1, 4, 8, 4, true>(n, out, arg, m, row0, rows, col0, stream);
            case -4: return lie_launch<T, 4, 1, 4, 4, 4, true>(n, out, arg, m, row0, rows, col0, stream);
            case -2: return lie_launch<T, 2, 1, 4, 2, 4, true>(n, out, arg, m, row0, rows, col0, stream);
            case -1: return lie_launch<T, 1, 1, 4, 1, 4, true>(n, out, arg, m, row0, rows, col0, stream);
            default: return fail(RPP_HIP_NOT_IMPLEMENTED, "Lie argument does not fit in LDS");
        }
    }
    switch (lie_objects_per_group<T>(m, arg.size) * (long_rows ? -1 : 1)) {
        case 8:
            return joint ? lie_launch<T, 8, 1, 2, 8>(n, out, arg, m, row0, rows, col0, stream)
                         : lie_launch<T, 8, RS, 1>(n, out, arg, m, row0, rows, col0, stream);
        case 4:
            return joint ? lie_launch<T, 4, 1, 2, 4>(n, out, arg, m, row0, rows, col0, stream)
                         : lie_launch<T, 4, RS, 1>(n, out, arg, m, row0, rows, col0, stream);
        case 2:
            return joint ? lie_launch<T, 2, 1, 2, 2>(n, out, arg, m, row0, rows, col0, stream)
                         : lie_launch<T, 2, RS, 1>(n, out, arg, m, row0, rows, col0, stream);
        case 1: return lie_launch<T, 1, RS, 1>(n, out, arg, m, row0, rows, col0, stream);
        case -8: return lie_launch<T, 8, 1, 4, 1, 8, true>(n, out, arg, m, row0, rows, col0, stream);
        case -4: return lie_launch<T, 4, 1, 4, 1, 8, true>(n, out, arg, m, row0, rows, col0, stream);
        case -2: return lie_launch<T, 2, 1, 4, 1, 8, true>(n, out, arg, m, row0, rows, col0, stream);
        case -1: return lie_launch<T, 1, 1, 4, 1, 8, true>(n, out, arg, m, row0, rows, col0, stream);
        default: return fail(RPP_HIP_NOT_IMPLEMENTED, "Lie argument does not fit in LDS");
    }
}

template <typename T>
bool lie_spmv_fits(const DeviceCsr& m, const rpp_hip_vector_batch& arg) {
    return m.table_size > 0 && m.rowdesc && m.packed4 && lie_objects_per_group<T>(m, arg.size) > 0;
}

template int launch_lie_spmv<float>(int64_t, const rpp_hip_vector_batch&, const rpp_hip_vector_batch&,
                                    const DeviceCsr&, int64_t, int64_t, int64_t, hipStream_t);
template int launch_lie_spmv<double>(int64_t, const rpp_hip_vector_batch&, const rpp_hip_vector_batch&,
                                     const DeviceCsr&, int64_t, int64_t, int64_t, hipStream_t);
template bool lie_spmv_fits<float>(const DeviceCsr&, const rpp_hip_vector_batch&);
template bool lie_spmv_fits<double>(const DeviceCsr&, const rpp_hip_vector_batch&);

}  // namespace rpp_hip
