// fv_scratch.hip -- the fixed device buffer behind fivonagi_step_noalloc.
//
// That entry point (fv_runtime.cpp) keeps everything a step needs beyond the
// caller's state -- the instance description, the two control blocks, the
// three sets of CFL reduction slots, the edge tables, the work items and the
// CFL lane masks -- in this buffer instead of allocating device memory.
//
// Every GPU the library runs on holds its own copy of the array, from the
// first use until the process ends.  In host code the array's name is not a
// device address: the address of a device's copy comes from
// hipGetSymbolAddress, looked up once per device (with that device current)
// and kept; only that address is handed to kernels and copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/fivonagi.h"

static __device__ __align__(256) unsigned char scratch[FIVONAGI_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];   // device address of the array, per device id
}  // namespace

// Address and size of the current device's copy of the array.
extern "C" hipError_t fv_scratch_get(void** ptr, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  *bytes = sizeof(scratch);
  return hipSuccess;
}
