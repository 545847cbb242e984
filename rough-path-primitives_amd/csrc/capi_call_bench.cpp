// Per-call cost of the C-ABI as a C++ caller (the rpp::hip drop-in layer)
// reaches it: rpp_hip_ft_mul_f64 on a batch of width-4, depth-4 pairs, called
// back to back on one stream with no other host work in between.
//
//   capi_call_bench [pairs] [calls]
//
// Prints one JSON line: the host time per call of the enqueueing loop
// (descriptor validation and the kernel launch), the wall time per call until
// the stream drains, and the device time per call between two events around
// the loop.  Inputs are zeros (the kernel's work does not depend on them).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "rpp_hip.h"

namespace {

int check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 1024;
    const int calls = argc > 2 ? std::atoi(argv[2]) : 2000;
    const int width = 4, depth = 4;
    const long long dim = 341;  // 1 + 4 + 16 + 64 + 256
    double *out = nullptr, *lhs = nullptr, *rhs = nullptr;
    const size_t bytes = static_cast<size_t>(n * dim) * sizeof(double);
    check(hipMalloc(&out, bytes), "hipMalloc");
    check(hipMalloc(&lhs, bytes), "hipMalloc");
    check(hipMalloc(&rhs, bytes), "hipMalloc");
    check(hipMemset(lhs, 0, bytes), "hipMemset");
    check(hipMemset(rhs, 0, bytes), "hipMemset");
    hipStream_t s;
    check(hipStreamCreate(&s), "hipStreamCreate");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    rpp_hip_ft_mul_desc d{};
    d.basis = rpp_hip_basis{width, depth};
    d.num_batches = n;
    d.out = rpp_hip_tensor_batch{out, dim, 1, 0, depth};
    d.lhs = rpp_hip_tensor_batch{lhs, dim, 1, 0, depth};
    d.rhs = rpp_hip_tensor_batch{rhs, dim, 1, 0, depth};
    d.beta = 1.0;
    for (int i = 0; i < 200; ++i) {
        if (int rc = rpp_hip_ft_mul_f64(&d, s)) {
            std::fprintf(stderr, "rpp_hip_ft_mul_f64: %d %s\n", rc, rpp_hip_last_error());
            return 1;
        }
    }
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
    const auto t0 = std::chrono::steady_clock::now();
    check(hipEventRecord(e0, s), "hipEventRecord");
    for (int i = 0; i < calls; ++i) rpp_hip_ft_mul_f64(&d, s);
    const auto t1 = std::chrono::steady_clock::now();
    check(hipEventRecord(e1, s), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    const auto t2 = std::chrono::steady_clock::now();
    float dev_ms = 0;
    check(hipEventElapsedTime(&dev_ms, e0, e1), "hipEventElapsedTime");
    const double host_us = std::chrono::duration<double, std::micro>(t1 - t0).count() / calls;
    const double wall_us = std::chrono::duration<double, std::micro>(t2 - t0).count() / calls;
    std::printf("{\"pairs\": %lld, \"calls\": %d, \"host_enqueue_us_per_call\": %.3f, \"wall_us_per_call\": %.3f, "
                "\"device_us_per_call\": %.3f}\n",
                n, calls, host_us, wall_us, 1e3 * dev_ms / calls);
    (void)hipFree(out);
    (void)hipFree(lhs);
    (void)hipFree(rhs);
    return 0;
}
