// Fixed device work array for the allocation-free entry point
// (sxd_gh_rhs_noalloc): a file-scope device array, one copy per GPU the
// library runs on, held until the process ends.
//
// In host code the array's name is a host-side stand-in, not a device
// address.  Its address on each device is looked up once per device with
// hipGetSymbolAddress (the device current at the time), kept per device, and
// only that address is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <mutex>

#include "kernels.hpp"

namespace sxd {
namespace {

// 16 Mi doubles = 128 MiB per device.
constexpr size_t kNoallocDoubles = size_t(16) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoallocDoubles];

std::mutex g_lookup_mutex;
std::atomic<double*> g_scratch_ptr[kMaxDevices];

}  // namespace

size_t noalloc_scratch_doubles() { return kNoallocDoubles; }

// The device address of `scratch` on the current device (looked up on the
// first call per device).  SXD_OK, or SXD_ERR_DEVICE if there is no current
// device or the symbol cannot be resolved.
int noalloc_scratch(double** ptr) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SXD_ERR_DEVICE;
  double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    std::lock_guard<std::mutex> lock(g_lookup_mutex);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (p == nullptr) {
      void* addr = nullptr;
      if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || addr == nullptr)
        return SXD_ERR_DEVICE;
      p = static_cast<double*>(addr);
      g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *ptr = p;
  return SXD_OK;
}

}  // namespace sxd
