// Adjoint of ft_log (operations/intermediate/ft_log.hpp:66-85).
//
// The reference evaluates log(1 + Y), Y = arg restricted to degrees >= 1, by a
// Horner scheme: out = 0; for d = N..1: out += c_d 1, out <- out (x) Y on
// degrees [0, N-d+1], with c_d = (-1)^(d+1) / d.  Writing P_d for out after
// the identity is added at step d (degrees 0..N-d) and Q_d = P_d (x) Y
// (degrees <= N-d+1), log = Q_1 and P_d = Q_{d+1} + c_d 1.  Reverse mode:
//     Qbar_1 = G;  for d = 1..N:
//         Ybar   += L_{P_d}^* Qbar_d      (degrees 1..N-d+1)
//         Qbar_{d+1} = R_Y^* Qbar_d       (degrees 0..N-d)
// The forward pass stores every P_d (sum_m start(m) coefficients per object,
// about dim / (width - 1)); each step is one batched product or adjoint
// product launch over all objects.
#include "common.hpp"
#include "ft_ops.hpp"

#include <algorithm>

namespace rpp_hip {

namespace {

constexpr int kLogSlot = 4;  // library workspace slot of this operation

template <typename T>
__global__ void k_log_unit(T* u) {
    if (threadIdx.x == 0 && blockIdx.x == 0) u[0] = T(1);
}

// dst[b * stride + i] = value for i < count (zero fill or a constant degree 0)
template <typename T>
__global__ void k_log_fill(T* dst, int64_t stride, int64_t count, T value, int64_t n) {
    const int64_t total = n * count;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = e / count;
        dst[b * stride + (e - b * count)] = value;
    }
}

// dst (contiguous rows of dim) = src zero-extended outside [lo, hi)
template <typename T>
__global__ void k_log_load(T* dst, int64_t dim, View<const T> src, int64_t lo, int64_t hi, int64_t n) {
    const int64_t total = n * dim;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = e / dim;
        const int64_t i = e - b * dim;
        dst[e] = (lo <= i && i < hi) ? src.obj(b)[i * src.cs] : T(0);
    }
}

// grad window [lo, hi): degree 0 (index 0) gets 0, the rest Ybar (0 when
// ybar is null)
template <typename T>
__global__ void k_log_store(View<T> grad, const T* ybar, int64_t dim, int64_t lo, int64_t hi, int64_t n) {
    const int64_t span = hi - lo;
    const int64_t total = n * span;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = e / span;
        const int64_t i = lo + (e - b * span);
        grad.obj(b)[i * grad.cs] = (i == 0 || ybar == nullptr) ? T(0) : ybar[b * dim + i];
    }
}

unsigned fill_grid(int64_t total) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65535LL * 4)));
}

}  // namespace

template <typename T>
int launch_log_adjoint(const DevBasis& B, int64_t n, const rpp_hip_tensor_batch& grad_arg,
                       const rpp_hip_tensor_batch& arg, const rpp_hip_tensor_batch& cot, hipStream_t stream) {
    if (n <= 0) return 0;
    const int N = B.depth;
    const int64_t dim = B.start[N + 1];
    if (grad_arg.min_degree > grad_arg.max_degree) return 0;
    if (N == 0) {
        // the log of a depth-0 tensor is 0: zero gradient on the window
        const int64_t lo = B.start[grad_arg.min_degree], hi = B.start[grad_arg.max_degree + 1];
        k_log_store<T><<<fill_grid(n * (hi - lo)), 256, 0, stream>>>(make_view<T>(grad_arg), nullptr, dim, lo, hi,
                                                                     n);
        return check_hip(hipGetLastError(), "k_log_store launch");
    }
    // P_d for d = 1..N at offset poff(d) of each object's row of PS words
    int64_t PS = 0;
    for (int d = 1; d <= N; ++d) PS += B.start[N - d + 1];
    auto poff = [&](int d) {
        int64_t o = 0;
        for (int e = 1; e < d; ++e) o += B.start[N - e + 1];
        return o;
    };
    const size_t unit_elems = 256 / sizeof(T);
    const size_t elems = unit_elems + static_cast<size_t>(n) * (PS + 3 * dim);
    void* ws = nullptr;
    if (int rc = workspace(kLogSlot, elems * sizeof(T), &ws)) return rc;
    T* unit = static_cast<T*>(ws);
    T* P = unit + unit_elems;
    T* Qb = P + n * PS;
    T* Pb = Qb + n * dim;
    T* Yb = Pb + n * dim;

    auto pview = [&](int d) {
        return rpp_hip_tensor_batch{P + poff(d), PS, 1, 0, N - d};
    };
    auto yview = [&](int K) {
        rpp_hip_tensor_batch y = arg;
        y.min_degree = std::max(arg.min_degree, 1);
        y.max_degree = std::min(arg.max_degree, K);
        return y;
    };
    auto cval = [](int d) { return (d % 2 == 0 ? -1.0 : 1.0) / d; };

    k_log_unit<T><<<1, 64, 0, stream>>>(unit);
    const rpp_hip_tensor_batch unit_view{unit, 0, 1, 0, 0};
    // P_N = c_N 1
    k_log_fill<T><<<fill_grid(n), 256, 0, stream>>>(P + poff(N), PS, 1, T(cval(N)), n);
    if (int rc = check_hip(hipGetLastError(), "k_log_fill launch")) return rc;
    for (int d = N - 1; d >= 1; --d) {
        const rpp_hip_tensor_batch y = yview(N - d);
        if (y.min_degree <= y.max_degree) {
            if (int rc = launch_product<T>(B, n, pview(d), pview(d + 1), y, &unit_view, cval(d), 1.0, stream))
                return rc;
        } else {
            // no Y degrees in range: P_d = c_d 1 on [0, N-d]
            k_log_fill<T><<<fill_grid(n * B.start[N - d + 1]), 256, 0, stream>>>(P + poff(d), PS,
                                                                                 B.start[N - d + 1], T(0), n);
            k_log_fill<T><<<fill_grid(n), 256, 0, stream>>>(P + poff(d), PS, 1, T(cval(d)), n);
            if (int rc = check_hip(hipGetLastError(), "k_log_fill launch")) return rc;
        }
    }
    // Qbar_1 = G (zero-extended), Ybar = 0
    const int64_t clo = cot.min_degree <= cot.max_degree ? B.start[cot.min_degree] : 0;
    const int64_t chi = cot.min_degree <= cot.max_degree ? B.start[cot.max_degree + 1] : 0;
    k_log_load<T><<<fill_grid(n * dim), 256, 0, stream>>>(Qb, dim, make_cview<T>(cot), clo, chi, n);
    k_log_fill<T><<<fill_grid(n * dim), 256, 0, stream>>>(Yb, dim, dim, T(0), n);
    if (int rc = check_hip(hipGetLastError(), "k_log_load launch")) return rc;
    for (int d = 1; d <= N; ++d) {
        const int K = N - d + 1;
        const rpp_hip_tensor_batch y = yview(K);
        const rpp_hip_tensor_batch qv{Qb, dim, 1, 0, K};
        const rpp_hip_tensor_batch yb{Yb, dim, 1, 1, K};
        if (int rc = launch_adjoint<T>(true, B, n, yb, pview(d), qv, stream, true)) return rc;
        if (d == N) break;
        const rpp_hip_tensor_batch pb{Pb, dim, 1, 0, N - d};
        if (y.min_degree <= y.max_degree) {
            if (int rc = launch_adjoint<T>(false, B, n, pb, y, qv, stream)) return rc;
        } else {
            k_log_fill<T><<<fill_grid(n * dim), 256, 0, stream>>>(Pb, dim, dim, T(0), n);
            if (int rc = check_hip(hipGetLastError(), "k_log_fill launch")) return rc;
        }
        std::swap(Qb, Pb);
    }
    const int64_t lo = B.start[grad_arg.min_degree], hi = B.start[grad_arg.max_degree + 1];
    k_log_store<T><<<fill_grid(n * (hi - lo)), 256, 0, stream>>>(make_view<T>(grad_arg), Yb, dim, lo, hi, n);
    return check_hip(hipGetLastError(), "k_log_store launch");
}

template int launch_log_adjoint<float>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                       const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&, hipStream_t);
template int launch_log_adjoint<double>(const DevBasis&, int64_t, const rpp_hip_tensor_batch&,
                                        const rpp_hip_tensor_batch&, const rpp_hip_tensor_batch&, hipStream_t);

}  // namespace rpp_hip
