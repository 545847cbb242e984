// Shared device/host definitions of librpp_hip: the tensor basis as kernel
// argument, strided batch views, small-divisor arithmetic and error state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <utility>

#include "rpp_hip.h"

namespace rpp_hip {

// Largest depth accepted.  Every level must also satisfy width^k < 2^31 so
// that in-level word indices fit an unsigned 32-bit register.
constexpr int kMaxDepth = 30;

// TensorBasis(width, depth) flattened into a kernel argument: level sizes
// width^k and degree offsets start_of_degree(k) (basis/tensor_basis.hpp
// :122-130, basis/graded_basis.hpp:13-98).  The magic number turns division
// of an in-level index n < 2^31 by the width into a multiply and a shift:
// with L = ceil(log2 width), magic = ceil(2^(32+L) / width) < 2^33, so
// n * magic < 2^64, and the rounding excess n * (magic - 2^(32+L)/width)
// / 2^(32+L) < 2^-(L+1) <= 1/(2 width) keeps floor(n / width) exact.
struct DevBasis {
    int32_t width;
    int32_t depth;
    uint64_t magic;
    uint32_t shift;
    uint32_t pow[kMaxDepth + 2];
    int64_t start[kMaxDepth + 2];

    __host__ __device__ __forceinline__ uint32_t div_width(uint32_t n) const {
        return static_cast<uint32_t>((static_cast<uint64_t>(n) * magic) >> shift);
    }
};

// Summation order of the reference's CPU product, which fixes its rounding.
// cpu::single_thread multiplies in tiles of power 5 (32-wide subtiles,
// 1024-entry tiles; cpu/single_thread/operations/basic/detail/
// ft_multiply.hpp:25-26).  Its Operation constructor (:114-145) fixes per
// width the tile letters T and the first right-tiled output degree F, and
// multiply_oneside (:1554-1598) forms an output degree k
//   - below F by contiguous_level_terms (:985-1027): rhs degree j ascending;
//   - at and above F by right_tiled_terms (:1405-1502), right blocks of
//     R = 2T letters (1 when T = 0): j = k (lhs[0] rhs_k), j = 0
//     (lhs_k rhs[0]), then j = R .. k-1 ascending (:1297-1336), then
//     j = R-1 .. 1 descending (:1338-1403);
//   - at width 1 by width_one_terms (:1126-1173): j descending.
// Every product and every sum is rounded on its own (x86-64 without FMA), so
// a kernel that sums in this order with contraction off agrees bit for bit.
constexpr int cpu_tile_letters(int w) {
    if (w > 32) return 0;
    int letters = 1;
    if (w > 1 && w < 32)
        for (int size = w; size * w <= 32; size *= w) ++letters;
    return letters;
}
constexpr int cpu_first_tiled_degree(int w) {
    if (w > 32) return 2;
    if (w == 1 || w == 32) return 3;
    int size = 1;
    for (int i = 0; i < cpu_tile_letters(w); ++i) size *= w;
    int first = 2 * cpu_tile_letters(w) + 1;
    for (long long level = static_cast<long long>(size) * size * w; level <= 1024; level *= w) ++first;
    return first;
}
constexpr int cpu_right_letters(int w) { return cpu_tile_letters(w) == 0 ? 1 : 2 * cpu_tile_letters(w); }

// The t-th term (t = 0 .. k) of output degree k of lhs (x) rhs: the rhs
// degree j it adds (a caller skips the degrees its windows exclude, and in
// the tiled order j = k when lhs[0] == 0 and j = 0 when rhs[0] == 0), given
// F = cpu_first_tiled_degree(width), R = cpu_right_letters(width) and
// whether the width is 1.  Loop-free, so that in unrolled loops it folds to a
// constant; kernels take F and R from constexpr variables.
constexpr int cpu_term_fr(int F, int R, bool width1, int k, int t) {
    if (width1) return k - t;
    if (k < F) return t;
    if (t == 0) return k;
    if (t == 1) return 0;
    if (t <= k - R + 1) return R + t - 2;
    return k + 1 - t;
}
constexpr int cpu_term(int w, int k, int t) {
    return cpu_term_fr(cpu_first_tiled_degree(w), cpu_right_letters(w), w == 1, k, t);
}

// The same order without the j = 0 term (t = 0 .. k-1), for output degree
// k >= 1 of a Horner step, whose rhs has no degree 0.
constexpr int cpu_term_no_rhs0_fr(int F, int R, bool width1, int k, int t) {
    if (width1) return k - t;
    if (k < F) return t + 1;
    return cpu_term_fr(F, R, false, k, t == 0 ? 0 : t + 1);
}
constexpr int cpu_term_no_rhs0(int w, int k, int t) {
    return cpu_term_no_rhs0_fr(cpu_first_tiled_degree(w), cpu_right_letters(w), w == 1, k, t);
}

// Calls f(std::integral_constant<int, I>{}) for I = 0 .. N-1 in order: inside
// f, decltype(i)::value is a constant expression, so index arithmetic built
// from it (degree offsets, word counts, a term order) is evaluated by the
// compiler's front end rather than left to loop unrolling.
// v, opaque to the optimiser (an empty asm statement that claims to change
// it): values derived from the result are computed where they are used, not
// hoisted out of a loop to occupy registers across it.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Compile-time width when W > 0, the basis' runtime width otherwise.
template <int W>
struct Letters {
    __device__ __forceinline__ static uint32_t div(const DevBasis& B, uint32_t n) {
        if constexpr (W > 0) {
            return n / static_cast<uint32_t>(W);
        } else {
            return B.div_width(n);
        }
    }
    __device__ __forceinline__ static uint32_t width(const DevBasis& B) {
        if constexpr (W > 0) {
            return static_cast<uint32_t>(W);
        } else {
            return static_cast<uint32_t>(B.width);
        }
    }
};

// sum_j lhs_{k-j}[w / width^j] * rhs_j[w mod width^j] for output word w of
// degree k over j in [jlo, jhi], in the reference's CPU order (cpu_term), each
// product and sum rounded on its own (contraction is off in this function's
// scope, which is what decides it, also where it is inlined).  lhs0_zero / rhs0_zero: lhs[0] / rhs[0] is zero (the tiled
// order skips those scalar terms).  lval(d, i) / rval(d, i) read word i of
// level d of lhs / rhs.  The split w -> (w / width^j, w mod width^j) advances
// one letter at a time as j rises, so the terms j = 1 .. R-1, which the
// tiled order adds last, are formed on the way up and held in registers.
// F and R of a compile-time width (W > 0), else of the basis' width.
template <int W>
struct CpuTiling {
    static constexpr int kF = W > 0 ? cpu_first_tiled_degree(W) : 0;
    static constexpr int kR = W > 0 ? cpu_right_letters(W) : 0;
    __device__ __forceinline__ static int first(const DevBasis& B) {
        if constexpr (W > 0) {
            return kF;
        } else {
            return cpu_first_tiled_degree(B.width);
        }
    }
    __device__ __forceinline__ static int right(const DevBasis& B) {
        if constexpr (W > 0) {
            return kR;
        } else {
            return cpu_right_letters(B.width);
        }
    }
};

// F and R: cpu_first_tiled_degree and cpu_right_letters of the width, which
// the caller evaluates once (constexpr for a compile-time width).
template <int W, typename T, typename Lf, typename Rf>
__device__ __forceinline__ T cpu_order_level_sum(const DevBasis& B, int F, int R, int k, uint32_t w, int jlo,
                                                 int jhi, bool lhs0_zero, bool rhs0_zero, Lf&& lval, Rf&& rval) {
#pragma clang fp contract(off)
    using Lt = Letters<W>;
    const uint32_t width = Lt::width(B);
    T acc = T(0);
    if (jlo > jhi) return acc;
    if (width == 1) {  // every level holds one word; j descending
        for (int j = jhi; j >= jlo; --j) acc += lval(k - j, 0u) * rval(j, 0u);
        return acc;
    }
    uint32_t p = w, s = 0;  // the split at j letters
    auto step = [&](int j) {  // j -> j + 1
        const uint32_t q = Lt::div(B, p);
        s += (p - q * width) * B.pow[j];
        p = q;
    };
    if (k < F) {
        for (int j = 0; j < jlo; ++j) step(j);
        for (int j = jlo;; ++j) {
            acc += lval(k - j, p) * rval(j, s);
            if (j == jhi) break;
            step(j);
        }
        return acc;
    }
    if (jlo <= k && k <= jhi && !lhs0_zero) acc += lval(0, 0u) * rval(k, w);
    if (jlo <= 0 && 0 <= jhi && !rhs0_zero) acc += lval(k, w) * rval(0, 0u);
    constexpr int kMaxLow = 9;  // R - 1 <= 9 (R = 10 at width 2)
    T low[kMaxLow];
#pragma unroll
    for (int q = 1; q <= kMaxLow; ++q) {
        low[q - 1] = T(0);
        if (q < R) {
            step(q - 1);
            if (q >= jlo && q <= jhi) low[q - 1] = lval(k - q, p) * rval(q, s);
        }
    }
    for (int j = R; j < k; ++j) {
        step(j - 1);
        if (j >= jlo && j <= jhi) acc += lval(k - j, p) * rval(j, s);
    }
#pragma unroll
    for (int q = kMaxLow; q >= 1; --q)
        if (q < R && q >= jlo && q <= jhi) acc += low[q - 1];
    return acc;
}

// One step of the reference's Horner recursion for ft_fmexp and ft_exp with a
// degree-one exponent (operations/intermediate/ft_fmexp.hpp:66-88,
// ft_exp.hpp:70-81): the product tile of a word is its single term added to
// the zero-initialised tile, 0 + a x; the writer scales it by c = 1/d and,
// on levels inside the multiplier's window, adds the multiplier's word s
// (ScaledTileAddWriter, ft_multiply.hpp:260-395), else writes c tile alone
// (ScaledTileWriter, :176-258).  Every operation rounds on its own, as on the
// reference's CPU, so chained steps agree with it bit for bit (signed zeros
// included).
template <typename T>
__device__ __forceinline__ T horner_ref(T a, T x, T c, T s, bool add) {
#pragma clang fp contract(off)
    const T tile = T(0) + a * x;
    const T v = c * tile;
    return add ? s + v : v;
}

// One strided batch of graded vectors (rpp_hip_tensor_batch on the device).
template <typename T>
struct View {
    T* data;
    int64_t bs;  // batch stride
    int64_t cs;  // coefficient stride
    int32_t mn;
    int32_t mx;

    __device__ __forceinline__ T* obj(int64_t b) const { return data + b * bs; }
    __device__ __forceinline__ bool has(int k) const { return mn <= k && k <= mx; }
};

// Streaming access for kernels that touch each coefficient once: nontemporal
// loads and stores mark the lines for early eviction, so a pass over a batch
// much larger than the caches does not displace lines still in use (a 1 GiB
// device copy runs 8% faster on MI355X with them).
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void stream_store(T v, T* p) {
    __builtin_nontemporal_store(v, p);
}

template <typename T>
View<T> make_view(const rpp_hip_tensor_batch& t) {
    return View<T>{static_cast<T*>(t.data), t.batch_stride, t.coeff_stride,
                   t.min_degree, t.max_degree};
}

template <typename T>
View<const T> make_cview(const rpp_hip_tensor_batch& t) {
    return View<const T>{static_cast<const T*>(t.data), t.batch_stride,
                         t.coeff_stride, t.min_degree, t.max_degree};
}

DevBasis make_basis(int width, int depth);

// Thread-local error message for rpp_hip_last_error().
void set_error(const std::string& msg);
const char* last_error_cstr();
int fail(int code, const std::string& msg);
int check_hip(hipError_t err, const char* what);

// Returns from the calling function with the code of a failed step.
#define RPP_TRY(expr)                     \
    do {                                  \
        if (int rc_ = (expr)) return rc_; \
    } while (0)

// Map a hipError_t to the rpp::ErrorCode numbering (cf. the reference's
// map_cuda_error, gpu/device.hpp:27-140).
int map_hip_error(hipError_t err);

// Grid size for a grid-stride loop over `items` work-groups.
unsigned grid_for(int64_t items, unsigned cap = 65535u * 8u);

// Library-owned device workspace of at least `bytes`, one per (device, slot),
// grown on demand and kept until unload.  Calls on one device that share a
// slot must be ordered on one stream.
int workspace(int slot, size_t bytes, void** ptr);

// Number of compute units of the current device (cached).
int device_cu_count();

}  // namespace rpp_hip
