// alpsg_disp_batch_noalloc: alpsg_disp_batch (det only) for callers that
// cannot allocate device memory -- code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc.  Its buffers beyond the inputs and
// outputs (the device copy of the omegas, the partial-chi slots, det D and the
// zero problem indices of the small-batch sweep) are carved from `scratch`, a
// file-scope device array of NOALLOC_BYTES that every GPU the library runs on
// holds until the process ends.  The host never passes the array's name to a
// kernel or copy: its address on each device comes from hipGetSymbolAddress,
// called once per device at the entry point's first call there.  Calls that
// overlap on one GPU share the array and race; callers serialise them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/alpsg.h"
#include "alpsg_guard.h"

// the disp chain on caller-given scratch (alpsg_api.cpp)
int alpsg_disp_on_scratch(alpsg_handle h, int n, const double *om, double *det,
                          unsigned char *scratch_dev, size_t scratch_bytes);

namespace {

constexpr size_t NOALLOC_BYTES = (size_t)32 << 20;  // 32 MiB per device
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_mu;
unsigned char *g_scratch_dev[MAX_DEVICES] = {nullptr};

// the device address of this device's copy of `scratch`
int scratch_address(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return alpsg_internal_fail(ALPSG_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return alpsg_internal_fail(ALPSG_E_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return alpsg_internal_fail(ALPSG_E_DEVICE,
                                 std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_dev[dev];
  return ALPSG_OK;
}

}  // namespace

extern "C" int alpsg_disp_batch_noalloc(alpsg_handle h, int n, const double *om, double *det) {
  ALPSG_TRY
  if (!h || (n > 0 && (!om || !det))) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  unsigned char *p = nullptr;
  const int rc = scratch_address(&p);
  if (rc) return rc;
  return alpsg_disp_on_scratch(h, n, om, det, p, NOALLOC_BYTES);
  ALPSG_CATCH
}
