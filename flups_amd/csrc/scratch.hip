// Static device scratch of flups_solve_noalloc: the stage arrays of a solve (the fused
// pipeline's S0 / S1, the general path's per-direction stage arrays) carved from one
// file-scope device array instead of memory the solver allocated, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
//
// Each GPU holds its own copy of `scratch` (kScratchBytes) from the load of the code object
// until the process ends. In host code the name `scratch` is a host-side stand-in, not a
// device address: the address of this device's copy comes from hipGetSymbolAddress, taken
// once per device at flups_setup with that device current, and only that pointer is handed
// to kernels. Every call on one GPU shares the array and finds the previous call's contents
// there; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "core.hpp"

namespace flups {

constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

// address of the current device's copy of `scratch`, taken once per device
void scratch_init() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    FLUPS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range for the static scratch", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev]) return;
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FLUPS_REQUIRE(p != nullptr && (reinterpret_cast<uintptr_t>(p) % 256) == 0, "static scratch address is not 256-byte aligned");
    g_ptr[dev] = static_cast<double*>(p);
}

// the current device's scratch (nullptr before scratch_init ran on it) and its size in bytes
double* scratch_device_ptr(size_t* bytes) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    *bytes = kScratchBytes;
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[dev];
}

}  // namespace flups
