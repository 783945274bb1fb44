// device.hpp -- device-side helpers: wave64 reductions, block reductions,
// precision conversion.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <type_traits>

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

// Four consecutive values from a 4-element-aligned address.
__device__ __forceinline__ void load4(const float* p, float (&v)[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
__device__ __forceinline__ void load4(const double* p, double (&v)[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}

// W consecutive elements at a 16-byte aligned address, moved as 16-byte
// words (W * sizeof(T) is a multiple of 16).
template <typename T, int W>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&v)[W]) {
  constexpr int NQ = W * (int)sizeof(T) / 16;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const uint4 u = reinterpret_cast<const uint4*>(p)[q];
    __builtin_memcpy(&v[q * 16 / (int)sizeof(T)], &u, 16);
  }
}
template <typename T, int W>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&v)[W]) {
  constexpr int NQ = W * (int)sizeof(T) / 16;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    uint4 u;
    __builtin_memcpy(&u, &v[q * 16 / (int)sizeof(T)], 16);
    reinterpret_cast<uint4*>(p)[q] = u;
  }
}

// W consecutive elements as one word of W * sizeof(T) = 4, 8 or 16 bytes
// (p aligned to that size).
template <typename T, int W>
__device__ __forceinline__ void wload(const T* __restrict__ p, T (&v)[W]) {
  constexpr int B = W * (int)sizeof(T);
  static_assert(B == 4 || B == 8 || B == 16, "word of 4, 8 or 16 bytes");
  if constexpr (B == 16) {
    vload<T, W>(p, v);
  } else if constexpr (B == 8) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    __builtin_memcpy(&v[0], &u, 8);
  } else {
    const unsigned u = *reinterpret_cast<const unsigned*>(p);
    __builtin_memcpy(&v[0], &u, 4);
  }
}
template <typename T, int W>
__device__ __forceinline__ void wstore(T* __restrict__ p, const T (&v)[W]) {
  constexpr int B = W * (int)sizeof(T);
  static_assert(B == 4 || B == 8 || B == 16, "word of 4, 8 or 16 bytes");
  if constexpr (B == 16) {
    vstore<T, W>(p, v);
  } else if constexpr (B == 8) {
    uint2 u;
    __builtin_memcpy(&u, &v[0], 8);
    *reinterpret_cast<uint2*>(p) = u;
  } else {
    unsigned u;
    __builtin_memcpy(&u, &v[0], 4);
    *reinterpret_cast<unsigned*>(p) = u;
  }
}

// CSR-stream staging: prod[k - c0] = val[k] * x[col[k]] for k in
// [c0, c0 + cn), by NT threads.  Each thread takes aligned groups of four
// entries (16-byte column and value loads; col and val must start on a
// 16-byte boundary, as device allocations do); entries before c0 in the
// first group are skipped, and the last, partial group is read entry by
// entry so nothing past c0 + cn is touched.
template <int NT, typename ACC, typename TV, typename TX>
__device__ __forceinline__ void stage_products(const int* __restrict__ col, const TV* __restrict__ val,
                                               const TX* __restrict__ x, int c0, int cn, ACC* prod) {
  const int end = c0 + cn;
  for (int idx = (c0 & ~3) + 4 * (int)threadIdx.x; idx < end; idx += 4 * NT) {
    if (idx + 3 < end) {
      const int4 c4 = *reinterpret_cast<const int4*>(col + idx);
      TV v[4];
      load4(val + idx, v);
      const int cc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (idx + e >= c0) prod[idx + e - c0] = (ACC)v[e] * (ACC)x[cc[e]];
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int k = idx + e;
        if (k >= c0 && k < end) prod[k - c0] = (ACC)val[k] * (ACC)x[col[k]];
      }
    }
  }
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

// 16-bit factor storage with a power-of-two scale s per scalar row: the
// stored word is the value times 1/s rounded to BF16 (uint16_t, round to
// nearest even) or FP16 (_Float16); scaling by a power of two is exact, so
// dequant(quant(v / s)) * s is v rounded to 16 bits relative to its row.
// bfloat16 from float, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Storage load/store: 16-bit entries carry the scale of their scalar row.
template <typename TC, typename TS>
__device__ __forceinline__ TC ldv(const TS* p, size_t i, float s) {
  if constexpr (std::is_same<TS, uint16_t>::value) {
    return __uint_as_float((uint32_t)p[i] << 16) * s;
  } else if constexpr (std::is_same<TS, _Float16>::value) {
    return (float)p[i] * s;
  } else {
    (void)s;
    return (TC)p[i];
  }
}
template <typename TC, typename TS>
__device__ __forceinline__ TS quant(TC v, float sinv) {
  if constexpr (std::is_same<TS, uint16_t>::value) {
    return f2bf(v * sinv);
  } else if constexpr (std::is_same<TS, _Float16>::value) {
    return (_Float16)(v * sinv);
  } else {
    (void)sinv;
    return (TS)v;
  }
}
template <typename TC, typename TS>
__device__ __forceinline__ TC dequant(TS q, float s) {
  if constexpr (std::is_same<TS, uint16_t>::value) {
    return __uint_as_float((uint32_t)q << 16) * s;
  } else if constexpr (std::is_same<TS, _Float16>::value) {
    return (float)q * s;
  } else {
    (void)s;
    return (TC)q;
  }
}

}  // namespace lace
