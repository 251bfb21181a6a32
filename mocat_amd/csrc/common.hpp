// Shared device-side building blocks for the mocat_amd kernels (gfx950, wave64).
//
//  * threefry2x32-20 and the counter layout jax.random uses for split/bits/uniform/normal
//    (the non-partitionable layout: a counter array of size S is cut into two halves,
//    x0 = counts[0:h], x1 = counts[h:S], padded with one zero when S is odd).
//    Call sites in the reference: mocat/src/ssm/filtering.py:181,267,278,294,
//    mocat/src/mcmc/standard_mcmc.py:94,120,129, mocat/src/mcmc/metropolis.py:60-61.
//  * uniform: bits >> 9 | 0x3F800000 reinterpreted as f32, minus one, then an affine map
//    evaluated with separately rounded multiply and add (no contraction).
//  * normal: sqrt(2) * erfinv(u) with u uniform on [nextafter(-1, 0), 1); erfinv is the
//    single-precision approximation published by M. Giles, "Approximating the erfinv
//    function", GPU Computing Gems Jade Edition (2011).
//  * wave / block reductions built on cross-lane shuffles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define MOCAT_DEV __device__ __forceinline__

namespace mocat {

constexpr int kWave = 64;

// ------------------------------------------------------------------------------------------
// threefry2x32 with 20 rounds (Salmon et al., SC'11). Rotation schedule and key-schedule
// parity constant are the published Threefry-2x32 values.
MOCAT_DEV uint32_t rotl32(uint32_t x, uint32_t r) { return (x << r) | (x >> (32u - r)); }

MOCAT_DEV void threefry2x32(uint32_t k0, uint32_t k1, uint32_t& x0, uint32_t& x1) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  x0 += k0;
  x1 += k1;
#define MOCAT_TF_ROUND(r) \
  x0 += x1;               \
  x1 = rotl32(x1, r);     \
  x1 ^= x0;
  MOCAT_TF_ROUND(13) MOCAT_TF_ROUND(15) MOCAT_TF_ROUND(26) MOCAT_TF_ROUND(6)
  x0 += k1; x1 += k2 + 1u;
  MOCAT_TF_ROUND(17) MOCAT_TF_ROUND(29) MOCAT_TF_ROUND(16) MOCAT_TF_ROUND(24)
  x0 += k2; x1 += k0 + 2u;
  MOCAT_TF_ROUND(13) MOCAT_TF_ROUND(15) MOCAT_TF_ROUND(26) MOCAT_TF_ROUND(6)
  x0 += k0; x1 += k1 + 3u;
  MOCAT_TF_ROUND(17) MOCAT_TF_ROUND(29) MOCAT_TF_ROUND(16) MOCAT_TF_ROUND(24)
  x0 += k1; x1 += k2 + 4u;
  MOCAT_TF_ROUND(13) MOCAT_TF_ROUND(15) MOCAT_TF_ROUND(26) MOCAT_TF_ROUND(6)
  x0 += k2; x1 += k0 + 5u;
#undef MOCAT_TF_ROUND
}

struct Key {
  uint32_t a, b;
};

// Element m of threefry_2x32(key, iota(S)) in the halves layout.
MOCAT_DEV uint32_t bits_at(Key k, uint32_t S, uint32_t m) {
  const uint32_t h = (S + 1u) >> 1;
  uint32_t x0, x1;
  if (m < h) {
    x0 = m;
    x1 = (m + h < S) ? m + h : 0u;
    threefry2x32(k.a, k.b, x0, x1);
    return x0;
  }
  x0 = m - h;
  x1 = m;
  threefry2x32(k.a, k.b, x0, x1);
  return x1;
}

// Both words of elements (i, i + h) come out of one threefry call: returns the pair for
// i < h. Used where a thread owns both halves (e.g. filling a whole bits array).
MOCAT_DEV void bits_pair(Key k, uint32_t S, uint32_t i, uint32_t& lo, uint32_t& hi) {
  const uint32_t h = (S + 1u) >> 1;
  uint32_t x0 = i, x1 = (i + h < S) ? i + h : 0u;
  threefry2x32(k.a, k.b, x0, x1);
  lo = x0;
  hi = x1;
}

// j-th key of split(key, num).
MOCAT_DEV Key split_at(Key k, uint32_t num, uint32_t j) {
  const uint32_t S = 2u * num;
  Key r;
  r.a = bits_at(k, S, 2u * j);
  r.b = bits_at(k, S, 2u * j + 1u);
  return r;
}

// f32 in [0, 1) from 32 random bits (23 mantissa bits kept).
MOCAT_DEV float unit_float(uint32_t bits) {
  return __uint_as_float((bits >> 9) | 0x3F800000u) - 1.0f;
}

// uniform(key, minval, maxval): max(minval, f * (maxval - minval) + minval), rounded step by step.
// HIP compiles with fp-contract=fast; the pragma keeps multiply and add separately rounded.
MOCAT_DEV float uniform_from_bits(uint32_t bits, float lo, float hi) {
#pragma clang fp contract(off)
  const float f = unit_float(bits);
  const float span = hi - lo;
  const float v = f * span + lo;
  return fmaxf(lo, v);
}

// Giles' single-precision erfinv approximation (central branch w < 5, tail branch w >= 5).
MOCAT_DEV float erfinv_f32(float x) {
  float w = -log1pf(-x * x);
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f;
    p = 3.43273939e-07f + p * w;
    p = -3.5233877e-06f + p * w;
    p = -4.39150654e-06f + p * w;
    p = 0.00021858087f + p * w;
    p = -0.00125372503f + p * w;
    p = -0.00417768164f + p * w;
    p = 0.246640727f + p * w;
    p = 1.50140941f + p * w;
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = 0.000100950558f + p * w;
    p = 0.00134934322f + p * w;
    p = -0.00367342844f + p * w;
    p = 0.00573950773f + p * w;
    p = -0.0076224613f + p * w;
    p = 0.00943887047f + p * w;
    p = 1.00167406f + p * w;
    p = 2.83297682f + p * w;
  }
  return (fabsf(x) == 1.0f) ? x * INFINITY : p * x;
}

// normal(key) element from its 32 random bits.
MOCAT_DEV float normal_from_bits(uint32_t bits) {
  const float lo = -0.99999994f;  // nextafter(-1, 0) in f32
  const float u = uniform_from_bits(bits, lo, 1.0f);
  return 1.41421354f * erfinv_f32(u);
}

// ------------------------------------------------------------------------------------------
// Reductions. Lane shuffles over the full 64-wide wavefront.
template <typename T>
MOCAT_DEV T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
template <typename T>
MOCAT_DEV T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}
template <typename T>
MOCAT_DEV T wave_min(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, kWave);
    v = w < v ? w : v;
  }
  return v;
}

// Copies a parameter struct (floats only, in parameter order) to shared memory with the whole
// workgroup, so that its global loads overlap; the caller synchronises before use.
template <class Params>
MOCAT_DEV void load_shared_params(Params& m, const float* p) {
  static_assert(sizeof(Params) == Params::NPARAM * sizeof(float), "parameter struct must be NPARAM floats");
  float* dst = reinterpret_cast<float*>(&m);
  for (int i = threadIdx.x; i < Params::NPARAM; i += blockDim.x) dst[i] = p[i];
}

// Block-wide reductions; `scratch` must hold blockDim.x / 64 elements. All threads get the result.
template <typename T>
MOCAT_DEV T block_sum(T v, T* scratch) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  T r = scratch[0];
  for (int i = 1; i < nw; ++i) r += scratch[i];
  return r;
}
template <typename T>
MOCAT_DEV T block_max(T v, T* scratch) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  T r = scratch[0];
  for (int i = 1; i < nw; ++i) r = scratch[i] > r ? scratch[i] : r;
  return r;
}

// Inclusive prefix sum within a wavefront (Hillis-Steele over shuffles).
template <typename T>
MOCAT_DEV T wave_inclusive_scan(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T w = __shfl_up(v, o, kWave);
    if (lane >= o) v += w;
  }
  return v;
}

}  // namespace mocat

// Status codes shared by every extern "C" entry point.
#define MOCAT_OK 0
#define MOCAT_ERR_ARG -1
#define MOCAT_ERR_HIP -2
#define MOCAT_ERR_SIZE -3

#define MOCAT_LAUNCH_CHECK()                               \
  do {                                                     \
    hipError_t e__ = hipGetLastError();                    \
    if (e__ != hipSuccess) return MOCAT_ERR_HIP;           \
  } while (0)
