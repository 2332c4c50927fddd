// ig_noalloc.hip -- the fixed device array behind the no-allocation J/K entry
// point (ig_direct_jk_device_noalloc): callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops) get their call
// scratch carved from this array instead of hipMalloc.  Each GPU the library
// runs on holds its own copy, at its own address, until the process ends;
// every call on one GPU shares it, so calls that overlap on one GPU race.
#include <mutex>

#include "ig_internal.h"

namespace ig {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(void **ptr, size_t *bytes) {
  static std::mutex m;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(m);
  if (!addr[dev]) {
    // the array's name is a host-side stand-in; the device address of this
    // GPU's copy comes from the runtime
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = kNoAllocScratchBytes;
  return hipSuccess;
}

}  // namespace ig
