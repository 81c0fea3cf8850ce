/*
 * Throughput probe for the multiply the mul/div/sqrt kernels are built on: each lane
 * runs 8 independent chains of 32x32->64 multiply-adds (v_mad_u64_u32).  The bench
 * divides by the measured rate (and also reports the quarter-rate figure derived from
 * the clock and CU count).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ __launch_bounds__(256) void ubench_mad(uint64_t *out, uint32_t iters, uint32_t seed) {
  uint64_t acc[8];
  uint32_t y = seed ^ threadIdx.x;
#pragma unroll
  for (int k = 0; k < 8; k++) acc[k] = (uint64_t)(threadIdx.x + k) * 0x9E3779B9u;
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * y + acc[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

/* Runs the probe with `blocks` x 256 threads and `iters` iterations (8 mads each) on the
 * null stream; returns the kernel time in milliseconds via *ms.  0 on success. */
extern "C" int fpa_ubench_mad(uint32_t blocks, uint32_t iters, float *ms) {
  uint64_t *out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * 8) != hipSuccess) return -1;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  hipLaunchKernelGGL(ubench_mad, dim3(blocks), dim3(256), 0, 0, out, iters / 8, 1u);
  (void)hipEventRecord(a, 0);
  hipLaunchKernelGGL(ubench_mad, dim3(blocks), dim3(256), 0, 0, out, iters, 1u);
  (void)hipEventRecord(b, 0);
  hipError_t e = hipEventSynchronize(b);
  float t = 0;
  (void)hipEventElapsedTime(&t, a, b);
  *ms = t;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  return e == hipSuccess ? (int)hipGetLastError() : (int)e;
}
