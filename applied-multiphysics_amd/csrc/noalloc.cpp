// Allocation-free solve (amp_solve_device_noalloc): the solve of amp_solve_device with its scratch (working-precision
// permuted vector, front vectors, backward partial products) taken from a fixed device array instead of the handle's
// growable buffers, for callers that must not allocate device memory.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "handles.h"

namespace {

constexpr size_t kScratchBytes = AMP_NOALLOC_SCRATCH_BYTES;

// One copy per device (each GPU the library runs on holds it until the process ends). In host code `scratch` names a
// host-side stand-in: kernels and copies only ever see the per-device address from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device with that device current.
int scratchAddress(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return amp::fail(AMP_ERR_HIP, "amp_solve_device_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return amp::fail(AMP_ERR_HIP, "amp_solve_device_noalloc: no device address for the scratch array");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return AMP_OK;
}

}  // namespace

extern "C" int amp_solve_device_noalloc(amp_num* num, const double* d_b, double* d_x, int32_t nrhs) {
  if (!num || nrhs < 0 || (nrhs > 0 && num->sym->S.n > 0 && (!d_b || !d_x)))
    return amp::fail(AMP_ERR_ARG, "amp_solve_device_noalloc: bad arguments");
  if (!num->ok) return amp::fail(AMP_ERR_STATE, "solve: no successful numeric factorization");
  if (num->sym->S.n == 0 || nrhs == 0) return AMP_OK;
  const amp::SolveScratch need = amp::solveScratchNeed(num, nrhs);
  if (need.total() > kScratchBytes)
    return amp::fail(AMP_ERR_ARG, "amp_solve_device_noalloc: " + std::to_string(nrhs) + " right-hand sides need " +
                                      std::to_string(need.total()) + " bytes of scratch, more than the " +
                                      std::to_string(kScratchBytes) + " the allocation-free solve holds");
  void* base = nullptr;
  if (int rc = scratchAddress(&base)) return rc;
  unsigned char* p = static_cast<unsigned char*>(base);
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return amp::solveWithScratch(num, d_b, d_x, nrhs, p, p + need.yBytes, p + need.yBytes + need.fvBytes);
}
