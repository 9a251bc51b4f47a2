// The fixed device scratch behind swap_assemble_static: one statically sized
// array per GPU, from which that entry point carves every device buffer of a
// build and merge, so that it calls no allocation routine.
//
// In host code the array's name is only a stand-in for its device address, so
// the address is looked up once per device (hipGetSymbolAddress, with that
// device current) when the first context is created on it, and only that
// pointer is ever handed to kernels or copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "swap_device.h"

namespace swapk {
// The largest input the static scratch takes is set from this size (see
// kStaticMaxInput in swap_common.h).
constexpr size_t kScratchBytes = kStaticScratchBytes;
}  // namespace swapk

static __device__ __align__(256) uint8_t scratch[swapk::kScratchBytes];

namespace swapk {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
}  // namespace

// Called with `device` current.  Idempotent per device; returns the error of
// the lookup, if it failed.
hipError_t static_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void *ptr = nullptr;
        g_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? ptr : nullptr;
        g_done[device] = true;
    }
    return g_err[device];
}

// Device address of the scratch on `device`, or null if set-up did not succeed.
void *static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_done[device] ? g_ptr[device] : nullptr;
}

}  // namespace swapk
