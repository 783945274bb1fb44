// device.hpp -- device-side helpers: wave64 reductions, block reductions,
// precision conversion.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

namespace lace {

// Full-wave (64 lanes) sum with DPP/permute shuffles.  Deterministic: the
// combination order is fixed by lane index.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over groups of G lanes (G a power of two <= 64); every lane of the
// group receives the group total.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum for blockDim.x = BS (multiple of 64); the result is valid in
// thread 0.  `smem` holds BS/64 elements.
template <int BS, typename T>
__device__ __forceinline__ T block_sum(T v, T* smem) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BS / 64; w++) r += smem[w];
  }
  __syncthreads();
  return r;
}

template <typename T>
struct Prec;
template <>
struct Prec<double> {
  __device__ __forceinline__ static double to_acc(double v) { return v; }
};
template <>
struct Prec<float> {
  __device__ __forceinline__ static float to_acc(float v) { return v; }
};

__device__ __forceinline__ double as_double(double v) { return v; }
__device__ __forceinline__ double as_double(float v) { return (double)v; }
__device__ __forceinline__ double as_double(__hip_bfloat16 v) { return (double)__bfloat162float(v); }

template <typename To, typename From>
__device__ __forceinline__ To convert(From v) {
  return static_cast<To>(v);
}

}  // namespace lace
