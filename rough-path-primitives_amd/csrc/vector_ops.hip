// Linear operations on batches of graded vectors (tensor or Lie):
// vector_assign, vector_add, vector_inplace_add, vector_scalar_multiply,
// vector_set_constant (operations/linalg/*.hpp), tensor_set_identity and
// tensor_add_identity (operations/basic/tensor_{set,add}_identity.hpp).
//
// A window is a contiguous range of coefficient indices, so the kernels take
// index ranges and never the basis.  All of them are HBM-bound streams:
//   - k_vector_map walks out's window with consecutive lanes on out's
//     unit-stride axis (coefficients for object-major batches, objects for
//     batch-interleaved ones) and reads the operands at the same (object,
//     coefficient), 0 outside their ranges;
//   - k_vector_transpose is vector_assign between an object-major and a
//     batch-interleaved batch: 64 x 64 (coefficient, object) tiles staged in
//     LDS, read along the source's unit-stride axis and written along the
//     destination's, so both sides move 512-byte runs; each work-group walks
//     a contiguous run of coefficient tiles with the next tile's loads in
//     flight during the current tile's stores.
// Contraction of a*b + c into one FMA is off in this file: every result is
// rounded as the reference's scalar loops round it
// (cpu/.../linalg/vector_add.hpp:52-60, vector_inplace_add.hpp:55-60), so
// the results agree with them bit for bit.
#include "common.hpp"
#include "vector_ops.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rpp_hip {

namespace {

struct DevRange {
    int64_t b;
    int64_t e;
    __device__ __forceinline__ bool has(int64_t i) const { return b <= i && i < e; }
};

template <typename T>
struct Operand {
    const T* p;
    int64_t bs;
    int64_t cs;
    DevRange r;
    __device__ __forceinline__ T at(int64_t b, int64_t i) const { return p[b * bs + i * cs]; }
};

template <int OP>
constexpr int operands_of() {
    return OP == kVecAdd ? 2 : (OP == kVecAssign || OP == kVecInplaceAdd) ? 1 : 0;
}

template <int OP>
constexpr bool reads_out() {
    return OP == kVecInplaceAdd || OP == kVecScale;
}

// One lane's value of out at coefficient i from the loaded operand values
// (x, y: 0 where the operand's range lacks i; o: out's old value).
template <typename T, int OP>
__device__ __forceinline__ T map_value(int64_t i, T o, T x, T y, bool hx, bool hy, T alpha, T beta) {
    if constexpr (OP == kVecAssign) {
        return x;
    } else if constexpr (OP == kVecAdd) {
        // Accum val{0}; val += alpha*lhs; val += beta*rhs (vector_add.hpp:52-60)
        T acc = T(0);
        if (hx) acc = acc + alpha * x;
        if (hy) acc = acc + beta * y;
        return acc;
    } else if constexpr (OP == kVecInplaceAdd) {
        return o + alpha * x;
    } else if constexpr (OP == kVecScale) {
        return o * alpha;
    } else if constexpr (OP == kVecSetConstant) {
        return alpha;
    } else {  // kVecSetIdentity
        return i == 0 ? alpha : T(0);
    }
}

constexpr int kMapUnroll = 4;

// out[b, i] for i in [ob, ob + W): the fast axis (consecutive lanes) is the
// coefficient when COEFF_FAST, the object otherwise.  A work-group is
// 2^lg_tx lanes along the fast axis times 256 / 2^lg_tx along the slow one,
// so short windows still fill whole wavefronts.  Each lane keeps
// kMapUnroll elements of its row in flight (all loads before any store), and
// the row's base addresses are formed once per row.  Stores are ordinary:
// non-temporal ones measured slower (vector_add over 4096 x 19531 FP64
// object-major, 366 -> 385 us).
template <typename T, int OP, bool COEFF_FAST>
__global__ void __launch_bounds__(256)
k_vector_map(T* out, int64_t obs, int64_t ocs, int64_t ob, int64_t W, int64_t n, int lg_tx, Operand<T> a,
             Operand<T> c, T alpha, T beta) {
    constexpr int NOP = operands_of<OP>();
    const int tx = threadIdx.x & ((1 << lg_tx) - 1);
    const int ty = threadIdx.x >> lg_tx;
    const int64_t TX = int64_t{1} << lg_tx;
    const int64_t TY = 256 >> lg_tx;
    const int64_t F = COEFF_FAST ? W : n;
    const int64_t S = COEFF_FAST ? n : W;
    // strides along the fast axis
    const int64_t of = COEFF_FAST ? ocs : obs;
    const int64_t af = COEFF_FAST ? a.cs : a.bs;
    const int64_t cf = COEFF_FAST ? c.cs : c.bs;
    for (int64_t s = blockIdx.y * TY + ty; s < S; s += gridDim.y * TY) {
        // the fast index f is at coefficient ob + f (COEFF_FAST) or object f
        T* orow = COEFF_FAST ? out + s * obs + ob * ocs : out + (ob + s) * ocs;
        const T* arow = nullptr;
        const T* crow = nullptr;
        if constexpr (NOP >= 1) arow = COEFF_FAST ? a.p + s * a.bs + ob * a.cs : a.p + (ob + s) * a.cs;
        if constexpr (NOP >= 2) crow = COEFF_FAST ? c.p + s * c.bs + ob * c.cs : c.p + (ob + s) * c.cs;
        for (int64_t f0 = blockIdx.x * TX * kMapUnroll + tx; f0 < F; f0 += gridDim.x * TX * kMapUnroll) {
            T x[kMapUnroll], y[kMapUnroll], o[kMapUnroll];
            bool hx[kMapUnroll], hy[kMapUnroll];
#pragma unroll
            for (int u = 0; u < kMapUnroll; ++u) {
                const int64_t f = f0 + u * TX;
                const int64_t i = ob + (COEFF_FAST ? f : s);
                const bool in = f < F;
                hx[u] = in && NOP >= 1 && a.r.has(i);
                hy[u] = in && NOP >= 2 && c.r.has(i);
                x[u] = T(0);
                y[u] = T(0);
                o[u] = T(0);
                if constexpr (NOP >= 1)
                    if (hx[u]) x[u] = arow[f * af];
                if constexpr (NOP >= 2)
                    if (hy[u]) y[u] = crow[f * cf];
                if constexpr (reads_out<OP>())
                    if (in) o[u] = orow[f * of];
            }
#pragma unroll
            for (int u = 0; u < kMapUnroll; ++u) {
                const int64_t f = f0 + u * TX;
                if (f < F)
                    orow[f * of] = map_value<T, OP>(ob + (COEFF_FAST ? f : s), o[u], x[u], y[u], hx[u], hy[u],
                                                    alpha, beta);
            }
        }
    }
}

// out[b, 0] += alpha for every object (tensor_add_identity.hpp:34-43).
template <typename T>
__global__ void __launch_bounds__(256) k_add_identity(T* out, int64_t obs, int64_t n, T alpha) {
    for (int64_t b = blockIdx.x * int64_t{256} + threadIdx.x; b < n; b += int64_t{gridDim.x} * 256) {
        T* o = out + b * obs;
        *o = *o + alpha;
    }
}

constexpr int kTile = 64;

// vector_assign across layouts.  OUT_INTERLEAVED: out is batch-interleaved
// (object stride 1) and arg object-major (coefficient stride 1); otherwise
// the reverse.  A tile holds tile[c][b] for 64 coefficients of out's window
// and 64 objects; the padding column keeps the transposed LDS accesses on
// distinct banks.  Work-group (x, y) owns object tile y (and every gridDim.y-th
// after it) and the contiguous run of coefficient tiles [x*chunk, (x+1)*chunk):
// consecutive tiles of one work-group write adjacent segments of the same
// object-major rows, whose partial cache lines then merge in L2, and the
// loads of the next tile are in flight while the current one is stored.
// Coefficients outside arg's range are 0 and never read.
template <typename T, bool OUT_INTERLEAVED>
__global__ void __launch_bounds__(256)
k_vector_transpose(T* out, int64_t obs, int64_t ocs, const T* arg, int64_t abs, int64_t acs, int64_t ob,
                   int64_t W, int64_t n, DevRange ar, int64_t chunk) {
    __shared__ T tile[kTile][kTile + 1];
    const int lane = threadIdx.x & (kTile - 1);
    const int row = threadIdx.x >> 6;  // 0..3
    const int64_t ctiles = (W + kTile - 1) / kTile;
    const int64_t t_begin = blockIdx.x * chunk;
    const int64_t t_end = min(t_begin + chunk, ctiles);
    T v[kTile / 4];
    for (int64_t b0 = blockIdx.y * int64_t{kTile}; b0 < n; b0 += gridDim.y * int64_t{kTile}) {
        // along arg's unit-stride axis: lanes walk coefficients of an
        // object-major arg, objects of an interleaved one
        auto load = [&](int64_t c0) {
#pragma unroll
            for (int r = 0; r < kTile / 4; ++r) {
                const int rr = row + 4 * r;
                const int cl = OUT_INTERLEAVED ? lane : rr;
                const int bl = OUT_INTERLEAVED ? rr : lane;
                const int64_t i = ob + c0 + cl;
                const int64_t b = b0 + bl;
                v[r] = (c0 + cl < W && b < n && ar.has(i)) ? arg[b * abs + i * acs] : T(0);
            }
        };
        if (t_begin < t_end) load(t_begin * kTile);
        for (int64_t t = t_begin; t < t_end; ++t) {
            const int64_t c0 = t * kTile;
#pragma unroll
            for (int r = 0; r < kTile / 4; ++r) {
                const int rr = row + 4 * r;
                tile[OUT_INTERLEAVED ? lane : rr][OUT_INTERLEAVED ? rr : lane] = v[r];
            }
            __syncthreads();
            if (t + 1 < t_end) load(c0 + kTile);
            // along out's unit-stride axis
#pragma unroll
            for (int r = 0; r < kTile / 4; ++r) {
                const int rr = row + 4 * r;
                const int cl = OUT_INTERLEAVED ? rr : lane;
                const int bl = OUT_INTERLEAVED ? lane : rr;
                const int64_t b = b0 + bl;
                if (c0 + cl < W && b < n) {
                    // non-temporal into an interleaved batch, whose rows are
                    // whole cache lines (4096 x 19531 FP64: 253 -> 233 us);
                    // ordinary stores into object-major rows, which start
                    // mid-line (288 us, against 338 non-temporal)
                    T* dst = &out[b * obs + (ob + c0 + cl) * ocs];
                    if constexpr (OUT_INTERLEAVED) {
                        stream_store(tile[cl][bl], dst);
                    } else {
                        *dst = tile[cl][bl];
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
Operand<T> operand_of(const rpp_hip_tensor_batch* t, IndexRange r) {
    if (!t) return Operand<T>{nullptr, 0, 1, DevRange{0, 0}};
    return Operand<T>{static_cast<const T*>(t->data), t->batch_stride, t->coeff_stride, DevRange{r.begin, r.end}};
}

template <typename T, int OP>
int launch_map(int64_t n, const rpp_hip_tensor_batch& out, IndexRange orr, Operand<T> a, Operand<T> c, T alpha,
               T beta, hipStream_t stream) {
    const int64_t W = orr.end - orr.begin;
    // object-fast only for a batch-interleaved out
    const bool coeff_fast = n == 1 || !(out.batch_stride == 1 && out.coeff_stride != 1);
    const int64_t F = coeff_fast ? W : n;
    const int64_t S = coeff_fast ? n : W;
    int lg_tx = 0;
    while (lg_tx < 8 && (int64_t{1} << lg_tx) < F) ++lg_tx;
    const int64_t TX = int64_t{1} << lg_tx, TY = 256 / TX;
    // about two work-groups per compute unit, each lane four elements deep:
    // the rows in flight at once then stay few and contiguous, which measured
    // fastest for these streams (5.2-5.3 TB/s at 4096 x 19531 FP64; more
    // resident work-groups spread the accesses over more DRAM pages)
    const int64_t target = std::max<int64_t>(1, int64_t{device_cu_count()} * 2);
    const int64_t gx = std::min<int64_t>((F + TX * kMapUnroll - 1) / (TX * kMapUnroll), target);
    const int64_t gy = std::min<int64_t>({(S + TY - 1) / TY, std::max<int64_t>(1, target / gx), 65535});
    const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
    T* o = static_cast<T*>(out.data);
    if (coeff_fast)
        k_vector_map<T, OP, true><<<grid, 256, 0, stream>>>(o, out.batch_stride, out.coeff_stride, orr.begin, W, n,
                                                             lg_tx, a, c, alpha, beta);
    else
        k_vector_map<T, OP, false><<<grid, 256, 0, stream>>>(o, out.batch_stride, out.coeff_stride, orr.begin, W, n,
                                                              lg_tx, a, c, alpha, beta);
    return check_hip(hipGetLastError(), "k_vector_map launch");
}

// The layout pair of an assign that the tiled transposition serves: 1 when
// out is object-major and arg batch-interleaved, 2 for the reverse, 0 else.
int transpose_kind(int64_t n, const rpp_hip_tensor_batch& out, const rpp_hip_tensor_batch& arg) {
    if (n < 2) return 0;
    const bool out_om = out.coeff_stride == 1 && out.batch_stride != 1;
    const bool out_il = out.batch_stride == 1 && out.coeff_stride != 1;
    const bool arg_om = arg.coeff_stride == 1 && arg.batch_stride != 1 && arg.batch_stride != 0;
    const bool arg_il = arg.batch_stride == 1 && arg.coeff_stride != 1;
    if (out_om && arg_il) return 1;
    if (out_il && arg_om) return 2;
    return 0;
}

}  // namespace

template <typename T>
int launch_vector_op(int op, int64_t n, const rpp_hip_tensor_batch& out, IndexRange out_range,
                     const rpp_hip_tensor_batch* a, IndexRange a_range, const rpp_hip_tensor_batch* c,
                     IndexRange c_range, double alpha, double beta, hipStream_t stream) {
    const int64_t W = out_range.end - out_range.begin;
    if (n <= 0 || W <= 0) return 0;
    const T al = static_cast<T>(alpha), be = static_cast<T>(beta);
    T* o = static_cast<T*>(out.data);
    if (op == kVecAddIdentity) {
        if (out_range.begin != 0) return 0;
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 65535));
        k_add_identity<T><<<grid, 256, 0, stream>>>(o, out.batch_stride, n, al);
        return check_hip(hipGetLastError(), "k_add_identity launch");
    }
    if (op == kVecAssign && a) {
        if (const int kind = transpose_kind(n, out, *a)) {
            // one work-group per compute unit, object tiles first, the
            // coefficient tiles split into as many runs as the units allow
            const int64_t cus = std::max(1, device_cu_count());
            const int64_t ctiles = (W + kTile - 1) / kTile, btiles = (n + kTile - 1) / kTile;
            const int64_t gy = std::min<int64_t>({btiles, cus, 65535});
            const int64_t runs = std::min<int64_t>(ctiles, std::max<int64_t>(1, cus / gy));
            const int64_t chunk = (ctiles + runs - 1) / runs;
            const dim3 grid(static_cast<unsigned>((ctiles + chunk - 1) / chunk), static_cast<unsigned>(gy));
            const DevRange ar{a_range.begin, a_range.end};
            const T* ap = static_cast<const T*>(a->data);
            if (kind == 2)
                k_vector_transpose<T, true><<<grid, 256, 0, stream>>>(o, out.batch_stride, out.coeff_stride, ap,
                                                                       a->batch_stride, a->coeff_stride,
                                                                       out_range.begin, W, n, ar, chunk);
            else
                k_vector_transpose<T, false><<<grid, 256, 0, stream>>>(o, out.batch_stride, out.coeff_stride, ap,
                                                                        a->batch_stride, a->coeff_stride,
                                                                        out_range.begin, W, n, ar, chunk);
            return check_hip(hipGetLastError(), "k_vector_transpose launch");
        }
    }
    const Operand<T> A = operand_of<T>(a, a_range), C = operand_of<T>(c, c_range);
    switch (op) {
        case kVecAssign: return launch_map<T, kVecAssign>(n, out, out_range, A, C, al, be, stream);
        case kVecAdd: return launch_map<T, kVecAdd>(n, out, out_range, A, C, al, be, stream);
        case kVecInplaceAdd: return launch_map<T, kVecInplaceAdd>(n, out, out_range, A, C, al, be, stream);
        case kVecScale: return launch_map<T, kVecScale>(n, out, out_range, A, C, al, be, stream);
        case kVecSetConstant: return launch_map<T, kVecSetConstant>(n, out, out_range, A, C, al, be, stream);
        case kVecSetIdentity: return launch_map<T, kVecSetIdentity>(n, out, out_range, A, C, al, be, stream);
        default: return fail(RPP_HIP_INTERNAL, "unknown vector operation");
    }
}

template int launch_vector_op<float>(int, int64_t, const rpp_hip_tensor_batch&, IndexRange,
                                     const rpp_hip_tensor_batch*, IndexRange, const rpp_hip_tensor_batch*,
                                     IndexRange, double, double, hipStream_t);
template int launch_vector_op<double>(int, int64_t, const rpp_hip_tensor_batch&, IndexRange,
                                      const rpp_hip_tensor_batch*, IndexRange, const rpp_hip_tensor_batch*,
                                      IndexRange, double, double, hipStream_t);

}  // namespace rpp_hip
