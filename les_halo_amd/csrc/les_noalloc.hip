// les_reduce_sum_exact_noalloc: the exact, order-independent FP32 sum of les_reduce_sum_exact
// for callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  It needs no handle: the per-workgroup partial limbs
// live in a file-scope device array of fixed size, and the result goes to a device float the
// caller passes, on the caller's stream.  Nothing here calls hipMalloc or hipFree.
#include <mutex>

#include "les_hip.h"
#include "les_kernels.h"

namespace les {
namespace {

// Partial limbs of at most NOALLOC_BLOCKS workgroups (160 KiB).  Each device holds its own copy
// of the array until the process ends; every call on a device shares it, so calls that overlap
// on one device race (serialise them, e.g. on one stream).
constexpr int NOALLOC_BLOCKS = 2048;
static __device__ __align__(256) int64_t scratch[NOALLOC_BLOCKS * NLIMB];

constexpr int MAX_DEVICES = 64;
std::once_flag g_once[MAX_DEVICES];
int64_t* g_scratch[MAX_DEVICES] = {};
hipError_t g_scratch_err[MAX_DEVICES] = {};

// The device address of this device's copy of 'scratch', looked up once per device with that
// device current (the array's name is only a host-side stand-in for it).
int64_t* scratch_of_current_device(int& dev) {
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::call_once(g_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_err[dev] == hipSuccess) g_scratch[dev] = static_cast<int64_t*>(p);
  });
  return g_scratch_err[dev] == hipSuccess ? g_scratch[dev] : nullptr;
}

}  // namespace
}  // namespace les

using namespace les;

extern "C" int les_reduce_sum_exact_noalloc(const float* dev_x, int64_t n, float* dev_out,
                                            void* stream) {
  if (!dev_x || !dev_out || n < 0 || n > LES_NOALLOC_SUM_MAX) return LES_E_INVALID;
  int dev = -1;
  int64_t* part = scratch_of_current_device(dev);
  if (!part) return LES_E_HIP;
  // every partial the finaliser reads is written by this call's first kernel: nothing in the
  // array has to start at zero
  const RedBuf rb{part, NOALLOC_BLOCKS};
  launch_exact_sum(static_cast<hipStream_t>(stream), dev_x, (long)n, rb, dev_out);
  return hipGetLastError() == hipSuccess ? LES_OK : LES_E_HIP;
}
