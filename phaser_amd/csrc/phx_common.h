// Common device/host helpers for the phaser MI355X (gfx950) backend.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PHX_OK 0
#define PHX_EINVAL (-1000)     // bad argument (shape, size, null pointer)
#define PHX_EUNSUPPORTED (-1001)  // size / combination without a kernel
#define PHX_ETOOLARGE (-1002)  // larger than a fixed-size device buffer of the library

// Return a negative hipError_t on launch failure.
#define PHX_CHECK_LAUNCH()                                   \
  do {                                                       \
    hipError_t e__ = hipGetLastError();                      \
    if (e__ != hipSuccess) return -(int)e__;                 \
  } while (0)

namespace phx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float abs2f(float2 a) { return a.x * a.x + a.y * a.y; }

// Hardware float atomics on coarse-grained device memory (global_atomic_add_f32/f64).
__device__ __forceinline__ void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_c64(float2* p, float2 v) {
  float* q = reinterpret_cast<float*>(p);
  unsafeAtomicAdd(q, v.x);
  unsafeAtomicAdd(q + 1, v.y);
}

// Sum of a double over the 64 lanes of a wave; every lane gets the result.
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of a double; every thread gets the result.  `scratch` holds
// blockDim.x/64 doubles of LDS.  All threads of the block must call it.
__device__ __forceinline__ double block_sum_f64(double v, double* scratch) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwave = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nwave; ++i) t += scratch[i];
  __syncthreads();
  return t;
}

__device__ __forceinline__ float block_max_f32(float v, float* scratch) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwave = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  float t = scratch[0];
  for (int i = 1; i < nwave; ++i) t = fmaxf(t, scratch[i]);
  __syncthreads();
  return t;
}

}  // namespace phx
