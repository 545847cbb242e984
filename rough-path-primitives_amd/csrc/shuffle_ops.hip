// The shuffle family: st_mul, st_fma, st_inplace_fma, st_adj_mul
// (operations/basic/st_*.hpp), left_hs_mul, left_hs_adj_lmul,
// left_hs_adj_rmul (operations/basic/left_hs_*.hpp).
//
// One thread per output coefficient (object, word).  Two sums cover the
// family:
//   - the shuffle sum of a word w of degree k: over the subsets S of its
//     positions, l[w|S] r[w|S^c], each factor only on its window; the left
//     half-shuffle fixes w's first letter on l's side
//     (operations/implementation/word_shuffle_product.hpp:49-93,
//     word_half_shuffle_product.hpp:20-66);
//   - the adjoint sum of an output word u of degree m: over op degrees j and
//     op words v of degree j, op[v] times the sum of arg over the C(m+j, j)
//     interleavings of u and v, optionally with one prefix letter on both
//     (word_shuffle_adjoint_multiply.hpp:22-95, shuffle_adjoint_op_loop.hpp
//     :20-60).
// Subsets are masks whose bit (L-1-p) marks position p (first letter first),
// visited in ascending order (basis/tensor_basis.hpp:58-91), and fixed-weight
// masks are walked by Gosper's successor, so every sum runs in the
// reference's order; with contraction of a*b + c off the results agree with
// it bit for bit.  Each thread keeps its word's letters in LDS, letter-major
// so the lanes of a wavefront read consecutive halves of dwords.
#include "common.hpp"
#include "shuffle_ops.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rpp_hip {

namespace {

constexpr int kShuffleBlock = 256;

template <typename T>
struct Operand {
    const T* obj;  // this object's coefficient 0
    int64_t cs;
    int mn, mx;
    __device__ __forceinline__ bool has(int k) const { return mn <= k && k <= mx; }
    __device__ __forceinline__ T at(int64_t i) const { return obj[i * cs]; }
};

struct Letters {
    uint16_t (*lt)[kShuffleBlock];
    int t;
    __device__ __forceinline__ uint32_t operator[](int p) const { return lt[p][t]; }
    __device__ __forceinline__ void set(int p, uint32_t v) const { lt[p][t] = static_cast<uint16_t>(v); }
};

// the letters of word `idx` of degree k, first letter at position off
__device__ __forceinline__ void unpack(const DevBasis& B, Letters L, int off, int k, uint32_t idx) {
    for (int p = off + k - 1; p >= off; --p) {
        const uint32_t q = B.div_width(idx);
        L.set(p, idx - q * static_cast<uint32_t>(B.width));
        idx = q;
    }
}

// sum over S of l[(pre) w|S] r[w|S^c] for the k letters of U from off;
// pre < 0: no prefix letter on l's side.
template <typename T>
__device__ T shuffle_sum(const DevBasis& B, Letters U, int off, int k, int pre, const Operand<T>& l,
                         const Operand<T>& r) {
    const int extra = pre >= 0 ? 1 : 0;
    const uint32_t W = static_cast<uint32_t>(B.width);
    T acc = T(0);
    for (uint32_t mask = 0; mask < (1u << k); ++mask) {
        const int ld = __popc(mask), rd = k - ld;
        if (!l.has(ld + extra) || !r.has(rd)) continue;
        uint32_t li = 0, ri = 0;
        for (int p = 0; p < k; ++p) {
            const uint32_t letter = U[off + p];
            if ((mask >> (k - 1 - p)) & 1u)
                li = li * W + letter;
            else
                ri = ri * W + letter;
        }
        const int64_t lidx = extra ? static_cast<int64_t>(pre) * B.pow[ld] + li : li;
        acc = acc + l.at(B.start[ld + extra] + lidx) * r.at(B.start[rd] + ri);
    }
    return acc;
}

// The adjoint sum of the m letters of U from uoff: op words carry the prefix
// op_pre and arg words the prefix arg_pre (negative: none).
template <typename T>
__device__ T adjoint_sum(const DevBasis& B, Letters U, Letters V, int uoff, int m, int op_pre, int arg_pre,
                         const Operand<T>& op, const Operand<T>& arg, int jlo, int jhi) {
    const int oe = op_pre >= 0 ? 1 : 0, ae = arg_pre >= 0 ? 1 : 0;
    const uint32_t W = static_cast<uint32_t>(B.width);
    T acc = T(0);
    for (int j = jlo; j <= jhi; ++j) {
        const int L = m + j;
        const int64_t obase = B.start[j + oe] + (oe ? static_cast<int64_t>(op_pre) * B.pow[j] : 0);
        const int64_t abase = B.start[L + ae] + (ae ? static_cast<int64_t>(arg_pre) * B.pow[L] : 0);
        for (uint32_t v = 0; v < B.pow[j]; ++v) {
            unpack(B, V, 0, j, v);
            const T opv = op.at(obase + v);
            T inner = T(0);
            uint32_t mask = (1u << j) - 1u;
            const uint32_t last = mask << m;
            for (;;) {
                int iu = uoff, iv = 0;
                uint32_t idx = 0;
                for (int p = 0; p < L; ++p) {
                    const uint32_t letter = ((mask >> (L - 1 - p)) & 1u) ? V[iv++] : U[iu++];
                    idx = idx * W + letter;
                }
                inner = inner + opv * arg.at(abase + idx);
                if (mask == last) break;
                const uint32_t lowest = mask & (0u - mask);
                const uint32_t ripple = mask + lowest;
                mask = ripple | (((ripple ^ mask) >> 2) / lowest);
            }
            acc = acc + inner;
        }
    }
    return acc;
}

template <typename T>
__device__ __forceinline__ Operand<T> operand(const View<const T>& v, int64_t b) {
    return Operand<T>{v.data ? v.obj(b) : nullptr, v.cs, v.mn, v.mx};
}

template <typename T, int OP>
__global__ void __launch_bounds__(kShuffleBlock)
k_shuffle(DevBasis B, int64_t n, View<T> out, View<const T> pv, View<const T> qv, View<const T> rv, T alpha,
          T beta) {
    __shared__ uint16_t ltu[kShuffleMaxDepth][kShuffleBlock];
    __shared__ uint16_t ltv[kShuffleMaxDepth][kShuffleBlock];
    const Letters U{ltu, static_cast<int>(threadIdx.x)}, V{ltv, static_cast<int>(threadIdx.x)};
    const int64_t ob = B.start[out.mn];
    const int64_t span = B.start[out.mx + 1] - ob;
    const int64_t total = n * span;
    for (int64_t item = blockIdx.x * int64_t{kShuffleBlock} + threadIdx.x; item < total;
         item += int64_t{gridDim.x} * kShuffleBlock) {
        const int64_t b = item / span;
        const int64_t i = ob + (item - b * span);
        int k = out.mn;
        while (k < out.mx && i >= B.start[k + 1]) ++k;
        const uint32_t local = static_cast<uint32_t>(i - B.start[k]);
        T* o = out.obj(b) + i * out.cs;
        const Operand<T> p = operand(pv, b), q = operand(qv, b);
        unpack(B, U, 0, k, local);
        T value;
        if constexpr (OP == kStMul || OP == kStFma || OP == kStInplace) {
            // st_mul.hpp:44-62, st_fma.hpp:55-76, st_inplace_fma.hpp:49-64
            if (k == 0) {
                const bool pq = p.has(0) && q.has(0);
                if constexpr (OP == kStMul) {
                    value = pq ? beta * p.at(0) * q.at(0) : T(0);
                } else if constexpr (OP == kStFma) {
                    const Operand<T> r = operand(rv, b);
                    value = T(0);
                    if (r.has(0)) value = value + alpha * r.at(0);
                    if (pq) value = value + beta * p.at(0) * q.at(0);
                } else {
                    value = alpha * *o;
                    if (pq) value = value + beta * p.at(0) * q.at(0);
                }
            } else {
                const T ws = shuffle_sum(B, U, 0, k, -1, p, q);
                if constexpr (OP == kStMul) {
                    value = beta * ws;
                } else if constexpr (OP == kStFma) {
                    const Operand<T> r = operand(rv, b);
                    value = T(0);
                    if (r.has(k)) value = value + alpha * r.at(i);
                    value = value + beta * ws;
                } else {
                    value = alpha * *o + beta * ws;
                }
            }
        } else if constexpr (OP == kStAdjMul) {
            // st_adj_mul.hpp:55-80: op degrees [max(op.min, arg.min - m), min(op.max, arg.max - m)]
            value = adjoint_sum(B, U, V, 0, k, -1, -1, p, q, max(p.mn, q.mn - k), min(p.mx, q.mx - k));
        } else if constexpr (OP == kHsMul) {
            // left_hs_mul.hpp:57-80: out[a w] = beta sum over S of x[a w|S] y[w|S^c]
            value = k == 0 ? T(0) : beta * shuffle_sum(B, U, 1, k - 1, static_cast<int>(U[0]), p, q);
        } else if constexpr (OP == kHsAdjL) {
            // left_hs_adj_lmul.hpp:88-116: sum over prefixes a of the adjoint sum with
            // op x[a v] and arg y[a merge], scaled by beta as the reference's GPU
            // operator scales it (gpu/block/operations/basic/left_hs_adj_lmul.hpp:114-177)
            const int jlo = max(0, max(p.mn - 1, q.mn - 1 - k)), jhi = min(p.mx - 1, q.mx - 1 - k);
            T acc = T(0);
            for (int a = 0; a < B.width; ++a) acc = acc + adjoint_sum(B, U, V, 0, k, a, a, p, q, jlo, jhi);
            value = beta * acc;
        } else {  // kHsAdjR
            // left_hs_adj_rmul.hpp:58-100: out[a u] = adjoint sum of u with op y[v] and arg x[a merge]
            value = k == 0 ? T(0)
                           : beta * adjoint_sum(B, U, V, 1, k - 1, -1, static_cast<int>(U[0]), q, p,
                                                max(q.mn, p.mn - k), min(q.mx, p.mx - k));
        }
        *o = value;
    }
}

template <typename T>
View<const T> cview_or_null(const rpp_hip_tensor_batch* t) {
    if (!t) return View<const T>{nullptr, 0, 1, 1, 0};
    return make_cview<T>(*t);
}

}  // namespace

template <typename T>
int launch_shuffle(int op, const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& out,
                   const rpp_hip_tensor_batch& p, const rpp_hip_tensor_batch& q, const rpp_hip_tensor_batch* r,
                   double alpha, double beta, hipStream_t stream) {
    if (n <= 0 || out.min_degree > out.max_degree) return 0;
    const int64_t span = B.start[out.max_degree + 1] - B.start[out.min_degree];
    const int64_t blocks = (n * span + kShuffleBlock - 1) / kShuffleBlock;
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks, device_cu_count() * 8)));
    const View<T> o = make_view<T>(out);
    const View<const T> P = make_cview<T>(p), Q = make_cview<T>(q), R = cview_or_null<T>(r);
    const T a = static_cast<T>(alpha), be = static_cast<T>(beta);
#define RPP_SHUFFLE(OP) k_shuffle<T, OP><<<grid, kShuffleBlock, 0, stream>>>(B, n, o, P, Q, R, a, be)
    switch (op) {
        case kStMul: RPP_SHUFFLE(kStMul); break;
        case kStFma: RPP_SHUFFLE(kStFma); break;
        case kStInplace: RPP_SHUFFLE(kStInplace); break;
        case kStAdjMul: RPP_SHUFFLE(kStAdjMul); break;
        case kHsMul: RPP_SHUFFLE(kHsMul); break;
        case kHsAdjL: RPP_SHUFFLE(kHsAdjL); break;
        case kHsAdjR: RPP_SHUFFLE(kHsAdjR); break;
        default: return fail(RPP_HIP_INTERNAL, "unknown shuffle operation");
    }
#undef RPP_SHUFFLE
    return check_hip(hipGetLastError(), "k_shuffle launch");
}

template int launch_shuffle<float>(int, const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                   const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,
                                   const rpp_hip_tensor_batch*, double, double, hipStream_t);
template int launch_shuffle<double>(int, const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                    const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&,
                                    const rpp_hip_tensor_batch*, double, double, hipStream_t);

}  // namespace rpp_hip
