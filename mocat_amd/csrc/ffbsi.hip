// Forward-filtering backward-simulation, full resampling (mocat/src/ssm/backward.py:20-39, 210-243).
//
// One backward step, for every trajectory m of M: key_m = split(key_t, M)[m] and
//   x_t[m] = particles_t[argmax_j(gumbel(key_m, (N,))[j] + log_weight_t[j] - U(particles_t[j] -> x_{t+1}[m]))]
// (random.categorical: the first maximum wins), U the model's transition potential as mocat's
// gaussian_potential evaluates it (utils.py:49-105): 0.5 |(x' - A x) @ S|^2 + nlz, S the transition
// precision square root. One wave per trajectory: the lanes take pairs (j, j + N/2), whose Gumbel bits
// come out of one threefry block (the halves layout), and keep a running maximum; a butterfly with
// the lower index winning ties combines them. Work per step: N x M potentials and N x M / 2 threefry
// blocks.
#include <cmath>

#include "common.hpp"

using namespace mocat;

namespace {

// linear Gaussian: params A[D*D], S[D*D] (transition precision square root), nlz
template <int D>
struct LgTrans {
  float A[D * D], S[D * D], nlz;
  MOCAT_DEV void load(const float* p) {
#pragma unroll
    for (int i = 0; i < D * D; ++i) A[i] = p[i];
#pragma unroll
    for (int i = 0; i < D * D; ++i) S[i] = p[D * D + i];
    nlz = p[2 * D * D];
  }
  MOCAT_DEV float pot(const float* xp, const float* xn) const {
    float diff[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) a += A[i * D + k] * xp[k];
      diff[i] = xn[i] - a;
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      float v = 0.f;
#pragma unroll
      for (int i = 0; i < D; ++i) v += diff[i] * S[i * D + k];
      s += v * v;
    }
    return 0.5f * s + nlz;
  }
};

// stochastic volatility, unit steps (stochastic_volatility.py:78-84): params phi, 1/sigma, nlz
struct SvTrans {
  float phi, isig, nlz;
  MOCAT_DEV void load(const float* p) {
    phi = p[0];
    isig = p[1];
    nlz = p[2];
  }
  MOCAT_DEV float pot(const float* xp, const float* xn) const {
    const float diff = xn[0] - xp[0] * phi;
    return 0.5f * (diff * diff) * (isig * isig) + nlz;
  }
};

constexpr float kTiny = 1.17549435e-38f;  // float32 tiny: the lower end of jax's gumbel uniforms

template <class Tr, int D>
__global__ __launch_bounds__(256) void ffbsi_step_kernel(const float* __restrict__ params,
                                                         const float* __restrict__ xs, const float* __restrict__ lw,
                                                         int64_t n, const float* __restrict__ xnext, int64_t m_count,
                                                         const uint32_t* __restrict__ key, int32_t* __restrict__ idx,
                                                         float* __restrict__ xout) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= m_count) return;
  Tr tr;
  tr.load(params);
  const Key km = split_at(Key{key[0], key[1]}, (uint32_t)m_count, (uint32_t)m);
  float xn[D];
#pragma unroll
  for (int q = 0; q < D; ++q) xn[q] = xnext[m * D + q];
  const uint32_t S = (uint32_t)n, h = (S + 1u) >> 1;
  float best = -INFINITY;
  int64_t bi = n;  // no candidate yet
  auto consider = [&](int64_t j, uint32_t bits) {
    const float u = uniform_from_bits(bits, kTiny, 1.f);
    const float v = -logf(-logf(u)) + (lw[j] - tr.pot(xs + j * D, xn));
    if (bi == n || v > best || (v == best && j < bi)) {
      best = v;
      bi = j;
    }
  };
  for (int64_t j = lane; j < h; j += kWave) {
    uint32_t w0, w1;
    bits_pair(km, S, (uint32_t)j, w0, w1);
    consider(j, w0);
    if (j + h < n) consider(j + h, w1);
  }
  // the maximum over the wave, the lower index on ties (first maximum)
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o, kWave);
    const int64_t oi = (int64_t)__shfl_xor((long long)bi, o, kWave);
    if (oi != n && (bi == n || ob > best || (ob == best && oi < bi))) {
      best = ob;
      bi = oi;
    }
  }
  if (bi == n) bi = 0;  // unreachable for n >= 1 (every lane pair holds a candidate after the butterfly)
  if (lane == 0 && idx) idx[m] = (int32_t)bi;
  if (lane < D) xout[m * D + lane] = xs[bi * D + lane];
}

template <class Tr, int D>
int launch_ffbsi(const float* params, const float* xs, const float* lw, int64_t n, const float* xnext, int64_t m,
                 const uint32_t* key, int32_t* idx, float* xout, hipStream_t stream) {
  const int per = 256 / kWave;
  hipLaunchKernelGGL((ffbsi_step_kernel<Tr, D>), dim3((unsigned)((m + per - 1) / per)), dim3(256), 0, stream, params,
                     xs, lw, n, xnext, m, key, idx, xout);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace

extern "C" {

int mocat_ffbsi_step(int model, int d, const float* params, const float* particles, const float* log_weight,
                     int64_t n, const float* x_next, int64_t m, const uint32_t* key, int32_t* idx, float* x_out,
                     hipStream_t stream) {
  if (!params || !particles || !log_weight || !x_next || !key || !x_out || n <= 0 || n >= 0xFFFFFFFFll || m <= 0 ||
      m >= 0xFFFFFFFFll)
    return MOCAT_ERR_ARG;
  if (model == 1) {
    if (d != 1) return MOCAT_ERR_ARG;
    return launch_ffbsi<SvTrans, 1>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
  }
  if (model != 0) return MOCAT_ERR_ARG;
  switch (d) {
    case 1: return launch_ffbsi<LgTrans<1>, 1>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 2: return launch_ffbsi<LgTrans<2>, 2>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 3: return launch_ffbsi<LgTrans<3>, 3>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 4: return launch_ffbsi<LgTrans<4>, 4>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 5: return launch_ffbsi<LgTrans<5>, 5>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    case 8: return launch_ffbsi<LgTrans<8>, 8>(params, particles, log_weight, n, x_next, m, key, idx, x_out, stream);
    default: return MOCAT_ERR_ARG;
  }
}

}  // extern "C"
