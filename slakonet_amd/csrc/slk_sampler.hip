// Live kernel sampling (see slk_sampler.h).
#include <mutex>
#include <vector>

#include "slakonet_amd.h"
#include "slk_sampler.h"

namespace {
constexpr size_t SLOT_ELEMS = (size_t)SLK_CLK_SUB * SLK_CLK_LINE;

__global__ void clk_init_kernel(unsigned long long* clk, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    clk[i] = (i % SLK_CLK_LINE) == 0 ? ~0ull : 0ull;
}

struct Sample {
  hipEvent_t e0, e1;
  double bytes, aux;
  int kind;
};
struct Sampler {
  int enabled = 0;
  int every = 0;
  long long launches[SLK_SAMPLE_KINDS] = {};
  double bytes[SLK_SAMPLE_KINDS] = {};
  double aux[SLK_SAMPLE_KINDS] = {};
  std::vector<Sample> pool;
  unsigned long long* clk = nullptr;   // device [capacity][2]
  size_t used = 0;
};
Sampler g;
std::mutex g_mu;
// byte hint of the next launch of each kind, per launching thread
thread_local double t_hint[SLK_SAMPLE_KINDS] = {};
thread_local double t_aux[SLK_SAMPLE_KINDS] = {};

void free_pool() {
  for (auto& s : g.pool) {
    (void)hipEventDestroy(s.e0);
    (void)hipEventDestroy(s.e1);
  }
  g.pool.clear();
  if (g.clk != nullptr) (void)hipFree(g.clk);
  g.clk = nullptr;
  g.used = 0;
}
}  // namespace

bool sampler_enabled() { return g.enabled; }

bool sampler_take(int kind, double bytes, SampleSlot* slot) {
  if (!g.enabled || kind < 0 || kind >= SLK_SAMPLE_KINDS) return false;
  double aux = 0.0;
  if (bytes < 0) {
    bytes = t_hint[kind];
    aux = t_aux[kind];
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g.enabled) return false;
  const long long idx = g.launches[kind]++;
  g.bytes[kind] += bytes;
  g.aux[kind] += aux;
  if (g.every <= 0 || idx % g.every != 0 || g.used >= g.pool.size()) return false;
  Sample& s = g.pool[g.used];
  s.bytes = bytes;
  s.aux = aux;
  s.kind = kind;
  slot->clk = g.clk + (size_t)g.used * SLOT_ELEMS;
  slot->e0 = s.e0;
  slot->e1 = s.e1;
  ++g.used;
  return true;
}

extern "C" {

int slk_sampler_reset(int enable, int every, int capacity) {
  std::lock_guard<std::mutex> lock(g_mu);
  (void)hipDeviceSynchronize();
  free_pool();
  for (int k = 0; k < SLK_SAMPLE_KINDS; ++k) {
    g.launches[k] = 0;
    g.bytes[k] = 0.0;
    g.aux[k] = 0.0;
  }
  g.enabled = enable;
  g.every = every;
  if (!enable || every <= 0 || capacity <= 0) return 0;
  const size_t elems = (size_t)capacity * SLOT_ELEMS;
  hipError_t e = hipMalloc((void**)&g.clk, elems * sizeof(unsigned long long));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(clk_init_kernel, dim3(1024), dim3(256), 0, 0, g.clk, elems);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    g.enabled = 0;
    return (int)e;
  }
  g.pool.resize((size_t)capacity);
  for (auto& s : g.pool) {
    // timing-only events: no system-scope fence around the kernel
    e = hipEventCreateWithFlags(&s.e0, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.e1, hipEventDisableSystemFence);
    if (e != hipSuccess) {
      g.enabled = 0;
      return (int)e;
    }
  }
  return 0;
}

int slk_sampler_hint(int kind, double bytes, double aux) {
  if (kind < 0 || kind >= SLK_SAMPLE_KINDS) return 1;
  t_hint[kind] = bytes;
  t_aux[kind] = aux;
  return 0;
}

int slk_sampler_read(int kind, double* out) {
  if (kind < 0 || kind >= SLK_SAMPLE_KINDS) return 1;
  std::lock_guard<std::mutex> lock(g_mu);
  for (int i = 0; i < 9; ++i) out[i] = 0.0;
  out[3] = (double)g.launches[kind];
  out[4] = g.bytes[kind];
  out[8] = g.aux[kind];
  if (g.used == 0) return 0;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return (int)e;
  std::vector<unsigned long long> clk(g.used * SLOT_ELEMS);
  e = hipMemcpy(clk.data(), g.clk, clk.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int)e;
  int dev = 0, khz = 0;
  (void)hipGetDevice(&dev);
  e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e != hipSuccess || khz <= 0) return e != hipSuccess ? (int)e : 1;
  double n = 0.0, ms = 0.0, bytes = 0.0, ev_ms = 0.0, aux = 0.0;
  for (size_t i = 0; i < g.used; ++i) {
    const Sample& s = g.pool[i];
    if (s.kind != kind) continue;
    float t = 0.0f;
    e = hipEventElapsedTime(&t, s.e0, s.e1);
    if (e != hipSuccess) return (int)e;
    unsigned long long t0 = ~0ull, t1 = 0ull;
    for (int sub = 0; sub < SLK_CLK_SUB; ++sub) {
      const unsigned long long* line = clk.data() + i * SLOT_ELEMS + (size_t)sub * SLK_CLK_LINE;
      if (line[0] < t0) t0 = line[0];
      if (line[1] > t1) t1 = line[1];
    }
    if (t1 < t0) continue;
    n += 1.0;
    ms += (double)(t1 - t0) / (double)khz;
    bytes += s.bytes;
    aux += s.aux;
    ev_ms += t;
  }
  out[0] = n;
  out[1] = ms;
  out[2] = bytes;
  out[5] = ev_ms;
  out[6] = (double)khz;
  out[7] = aux;
  return 0;
}

}  // extern "C"
