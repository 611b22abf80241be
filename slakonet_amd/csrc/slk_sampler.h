// Live kernel sampling. While enabled, every launch of a sampled kernel kind
// is counted with its algorithmic byte count, and every `every`-th launch is
// given a device slot in which its workgroups record the kernel's execution
// window: the earliest workgroup start (atomic min) and the latest
// workgroup end (atomic max) of the constant-rate wall clock. The same
// launch is also bracketed by a HIP start/stop event pair
// (hipExtLaunchKernelGGL) on its launch stream. Read back per kind after the
// work completes. Thread-safe (launch threads may call concurrently; the
// byte hint is per thread); off unless enabled.
#pragma once
#include <hip/hip_runtime.h>

// Kinds and their algorithmic amount: bytes for the memory-bound kinds (0, 1),
// flops for the FP64 MFMA GEMM (2), the bulge chase (3) and the FP32 fused
// stage-1 update (4) of complex Hermitian batches; bytes for the DOS (5),
// whose auxiliary amount is the number of Gaussians of the full sum
// (nk * nb * n_E per structure); flops of the real-symmetric instantiations of
// the fused update (6) and the chase (7), a quarter of the complex counts (one
// real multiply-add in place of four).
enum {
  SLK_SAMPLE_HETRD_STEP = 0,
  SLK_SAMPLE_BUILD = 1,
  SLK_SAMPLE_GEMM_F64 = 2,
  SLK_SAMPLE_CHASE = 3,
  SLK_SAMPLE_FUSED_AV = 4,
  SLK_SAMPLE_DOS = 5,
  SLK_SAMPLE_FUSED_AV_REAL = 6,
  SLK_SAMPLE_CHASE_REAL = 7,
  SLK_SAMPLE_KINDS = 8
};

// Workgroups of a sampled launch spread their clock atomics over
// SLK_CLK_SUB sub-slots (one 128-byte line each) to avoid contention on one
// address; the reader takes min / max over the sub-slots.
#define SLK_CLK_SUB 64
#define SLK_CLK_LINE 16   // unsigned long long per sub-slot line

struct SampleSlot {
  unsigned long long* clk;   // device [SLK_CLK_SUB][SLK_CLK_LINE]: [0] start (min), [1] end (max)
  hipEvent_t e0, e1;
};

// Counts one launch of `kind` moving `bytes` algorithmic bytes (bytes < 0:
// the pending hint of that kind, with its auxiliary amount). Returns true and
// a slot when this launch is sampled.
bool sampler_take(int kind, double bytes, SampleSlot* slot);
// Cheap check before computing an expensive algorithmic amount.
bool sampler_enabled();

// Device side: call at the start and at the end of every workgroup of a
// sampled launch (clk != nullptr); the end call must follow a barrier.
__device__ __forceinline__ unsigned long long* sample_line(unsigned long long* clk) {
  const unsigned wg = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  return clk + (size_t)(wg % SLK_CLK_SUB) * SLK_CLK_LINE;
}
__device__ __forceinline__ void sample_start(unsigned long long* clk) {
  if (clk != nullptr && threadIdx.x == 0 && threadIdx.y == 0)
    atomicMin(sample_line(clk), (unsigned long long)wall_clock64());
}
__device__ __forceinline__ void sample_end(unsigned long long* clk) {
  if (clk != nullptr && threadIdx.x == 0 && threadIdx.y == 0)
    atomicMax(sample_line(clk) + 1, (unsigned long long)wall_clock64());
}
