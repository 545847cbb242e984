// Shared device/host definitions of librpp_hip: the tensor basis as kernel
// argument, strided batch views, small-divisor arithmetic and error state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

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
