// jax.random-compatible counter RNG on the GPU: split / bits / uniform / normal, batched over
// keys. Each thread evaluates one threefry2x32 block and writes the two words it yields
// (elements i and i + h of the halves layout), so every key's output is contiguous.
#include "common.hpp"

using namespace mocat;

namespace {

enum class Out { Bits, Uniform, Normal };

template <Out kind, typename T>
__global__ __launch_bounds__(256) void rng_fill_kernel(const uint32_t* __restrict__ keys, int64_t n_keys,
                                                       uint32_t S, float lo, float hi, T* __restrict__ out) {
  const uint32_t h = (S + 1u) >> 1;
  const int64_t total = n_keys * (int64_t)h;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kid = g / h;
    const uint32_t i = (uint32_t)(g - kid * h);
    Key k{keys[2 * kid], keys[2 * kid + 1]};
    uint32_t w0, w1;
    bits_pair(k, S, i, w0, w1);
    T* o = out + kid * (int64_t)S;
    if constexpr (kind == Out::Bits) {
      o[i] = w0;
      if (i + h < S) o[i + h] = w1;
    } else if constexpr (kind == Out::Uniform) {
      o[i] = uniform_from_bits(w0, lo, hi);
      if (i + h < S) o[i + h] = uniform_from_bits(w1, lo, hi);
    } else {
      o[i] = normal_from_bits(w0);
      if (i + h < S) o[i + h] = normal_from_bits(w1);
    }
  }
}

// jax.random.randint(key, shape, minval, maxval) for int32 (the legacy algorithm): k1, k2 =
// split(key); hi, lo = bits(k1), bits(k2) of the request's size; span = maxval - minval as uint32 (1
// when maxval <= minval); m = (2^16 mod span)^2 mod span (the square wraps at 32 bits, as the uint32
// product does); result = minval + ((hi mod span) * m + (lo mod span)) mod span, with the products
// and sum wrapping at 32 bits.
MOCAT_DEV int32_t randint_from_bits(uint32_t hb, uint32_t lb, int32_t minval, int32_t maxval) {
  const uint32_t span = maxval <= minval ? 1u : (uint32_t)maxval - (uint32_t)minval;
  uint32_t m = 65536u % span;
  m = (m * m) % span;
  const uint32_t off = ((hb % span) * m + lb % span) % span;
  return (int32_t)((uint32_t)minval + off);
}

__global__ __launch_bounds__(256) void randint_kernel(const uint32_t* __restrict__ keys, int64_t n_keys, uint32_t S,
                                                      int32_t minval, int32_t maxval, int32_t* __restrict__ out) {
  const uint32_t h = (S + 1u) >> 1;
  const int64_t total = n_keys * (int64_t)h;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kid = g / h;
    const uint32_t i = (uint32_t)(g - kid * h);
    const Key k{keys[2 * kid], keys[2 * kid + 1]};
    uint32_t h0, h1, l0, l1;
    bits_pair(split_at(k, 2, 0), S, i, h0, h1);
    bits_pair(split_at(k, 2, 1), S, i, l0, l1);
    int32_t* o = out + kid * (int64_t)S;
    o[i] = randint_from_bits(h0, l0, minval, maxval);
    if (i + h < S) o[i + h] = randint_from_bits(h1, l1, minval, maxval);
  }
}

template <Out kind, typename T>
int launch_fill(const uint32_t* keys, int64_t n_keys, int64_t size, float lo, float hi, T* out,
                hipStream_t stream) {
  if (n_keys < 0 || size < 0 || (n_keys > 0 && (!keys || !out))) return MOCAT_ERR_ARG;
  // jax.random refuses requests whose counter does not fit 32 bits; so do we.
  if (size >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  if (n_keys == 0 || size == 0) return MOCAT_OK;
  const uint32_t S = (uint32_t)size;
  const int64_t total = n_keys * (int64_t)((S + 1u) >> 1);
  const int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((rng_fill_kernel<kind, T>), dim3((unsigned)blocks), dim3(threads), 0, stream, keys,
                     n_keys, S, lo, hi, out);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace

extern "C" {

int mocat_abi_version(void) { return 5; }

int mocat_random_split(const uint32_t* keys, int64_t n_keys, int64_t num, uint32_t* out, hipStream_t stream) {
  if (num < 0) return MOCAT_ERR_ARG;
  if (2 * num >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  return launch_fill<Out::Bits, uint32_t>(keys, n_keys, 2 * num, 0.f, 0.f, out, stream);
}

int mocat_random_bits(const uint32_t* keys, int64_t n_keys, int64_t size, uint32_t* out, hipStream_t stream) {
  return launch_fill<Out::Bits, uint32_t>(keys, n_keys, size, 0.f, 0.f, out, stream);
}

int mocat_random_uniform(const uint32_t* keys, int64_t n_keys, int64_t size, float minval, float maxval,
                         float* out, hipStream_t stream) {
  return launch_fill<Out::Uniform, float>(keys, n_keys, size, minval, maxval, out, stream);
}

int mocat_random_normal(const uint32_t* keys, int64_t n_keys, int64_t size, float* out, hipStream_t stream) {
  return launch_fill<Out::Normal, float>(keys, n_keys, size, 0.f, 0.f, out, stream);
}

int mocat_random_randint(const uint32_t* keys, int64_t n_keys, int64_t size, int32_t minval, int32_t maxval,
                         int32_t* out, hipStream_t stream) {
  if (n_keys < 0 || size < 0 || (n_keys > 0 && (!keys || !out))) return MOCAT_ERR_ARG;
  if (size >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  if (n_keys == 0 || size == 0) return MOCAT_OK;
  const uint32_t S = (uint32_t)size;
  const int64_t total = n_keys * (int64_t)((S + 1u) >> 1);
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(randint_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, keys, n_keys, S, minval, maxval,
                     out);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
