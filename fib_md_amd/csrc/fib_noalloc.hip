// fib_noalloc.hip -- fib_force_noalloc: a force evaluation that allocates no
// device memory, for callers that cannot call hipMalloc (code under HIP
// graph capture, tight serving loops, harnesses that forbid it).
//
// Every intermediate array (sorted copies, cell buckets, neighbour lists,
// leg forces, reduction partials, the pair table) lives in `scratch`, a
// file-scope device array of FIB_NOALLOC_BYTES that each GPU the library
// runs on holds until the process ends.  Its device address is looked up
// once per device with hipGetSymbolAddress at the first call on that device;
// host code uses that address, never the array's name.  All calls on one
// GPU share the array: calls that overlap on one GPU race, and every call
// rewrites what it uses (nothing carries over from the previous call).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/fibgpu.h"
#include "fib_kernels.h"

static __device__ __align__(256) unsigned char scratch[FIB_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

extern "C" int32_t fib_force_noalloc(const fib_params *p, int32_t precision, int32_t natm,
                                     int32_t nsg, int32_t nlj, const double lb[3],
                                     int32_t impact, const double *X, const int32_t *atype,
                                     double *F, double *utot, void *stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FIB_ERR_NODEVICE;
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) != hipSuccess ||
          std::string(prop.gcnArchName).compare(0, 6, "gfx950") != 0)
        return FIB_ERR_NODEVICE;
      void *a = nullptr;
      if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a) return FIB_ERR_HIP;
      g_scratch_ptr[dev] = a;
    }
    ptr = g_scratch_ptr[dev];
  }
  return fib_arena_force(p, precision, natm, nsg, nlj, lb, impact, X, atype, F, utot, ptr,
                         FIB_NOALLOC_BYTES, FIB_NOALLOC_MAX_ATOMS,
                         static_cast<hipStream_t>(stream));
}
