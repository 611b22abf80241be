// Allocation-free eigensolve: slk_eigh_mixed with its workspace in a static
// device array instead of a caller buffer, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The array `scratch` is a file-scope device symbol; in host code its name is
// only a host-side stand-in, so its device address is looked up once per
// device with hipGetSymbolAddress (each GPU holds its own copy of the array)
// and only that address is ever passed to kernels. Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU (different streams,
// or different host threads) race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "slakonet_amd.h"

#define SLK_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define SLK_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[SLK_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_mu;
void* g_ptr[SLK_NOALLOC_MAX_DEVICES] = {};

// Device address of this GPU's copy of `scratch` (the current device's),
// looked up on the first call on that device.
hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= SLK_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}
}  // namespace

extern "C" {

size_t slk_noalloc_scratch_bytes(void) { return (size_t)SLK_NOALLOC_SCRATCH_BYTES; }

int slk_eigh_mixed_noalloc(const void* H, const void* S, int n, int ld, int batch, double* w, int wld, int* info,
                           hipStream_t stream) {
  const size_t need = slk_eigh_mixed_workspace_bytes(n, ld, batch);
  if (need > (size_t)SLK_NOALLOC_SCRATCH_BYTES) return 4;
  void* ws = nullptr;
  const hipError_t e = scratch_address(&ws);
  if (e != hipSuccess) return (int)e;
  return slk_eigh_mixed(H, S, n, ld, batch, w, wld, ws, need, info, stream);
}

}  // extern "C"
