// Per-device runtime state of librpp_hip: grow-only workspaces and device
// properties, created on first use and kept until unload.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "common.hpp"

namespace rpp_hip {

namespace {
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, int>, Buffer> g_ws;
}  // namespace

int workspace(int slot, size_t bytes, void** ptr) {
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    std::lock_guard<std::mutex> lock(g_ws_mu);
    Buffer& b = g_ws[{dev, slot}];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // Work queued on other streams may still use the old buffer.
            if (int rc = check_hip(hipDeviceSynchronize(), "synchronising before growing a workspace")) return rc;
            if (int rc = check_hip(hipFree(b.ptr), "hipFree")) return rc;
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t grow = bytes + bytes / 4;
        if (hipMalloc(&b.ptr, grow) != hipSuccess) {
            b.ptr = nullptr;
            return fail(RPP_HIP_OUT_OF_RESOURCES, "workspace allocation failed");
        }
        b.bytes = grow;
    }
    *ptr = b.ptr;
    return 0;
}

int device_cu_count() {
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    cache[dev] = cus;
    return cus;
}

}  // namespace rpp_hip
