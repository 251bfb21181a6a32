// The linear-Gaussian model's per-particle functions at a runtime dimension, for the (d, d_y) outside the
// shapes pf.hip compiles: the bootstrap filter's initial sample and weight, proposal and weight, and the
// likelihood potential (mocat/src/ssm/linear_gaussian/linear_gaussian.py:46-51, 86-94, 118-128;
// mocat/src/ssm/filtering.py:65-78, 142-170), d and d_y up to kLgMaxDim. One thread per particle, its
// vectors in LDS (rows indexed at run time), the parameters in LDS, the formulas and summation order of
// pf.hip's Lgssm<D, DY>, and the noise drawn from the particle's key as normal(key, (d,)) in jax's
// layout. mocat_pf_hook sends model 0 here when its (d, d_y) is not compiled; the host runs such a
// filter's time loop over these hooks and the library's resampling, gather and log-sum-exp kernels.
#include "common.hpp"

namespace mocat {

constexpr int kLgMaxDim = 16;

namespace {

constexpr int kLgThreads = 128;  // per-thread vectors: 3 x 16 x 128 floats = 24 KB of LDS per workgroup
constexpr int kLgMaxParams = kLgMaxDim + 5 * kLgMaxDim * kLgMaxDim + 1;

struct LgView {
  const float *m0, *L0, *A, *LQ, *H, *SR;
  float nlz;
  int d, dy;
};

// 0.5 |(y - H s) @ SR|^2 + nlz for the thread's vector s (column t of S); T receives y - H s.
MOCAT_DEV float lg_lik(const LgView& m, const float (*S)[kLgThreads], float (*T)[kLgThreads], const float* y,
                       int t) {
  for (int j = 0; j < m.dy; ++j) {
    float hx = 0.f;
    for (int k = 0; k < m.d; ++k) hx += m.H[j * m.d + k] * S[k][t];
    T[j][t] = y[j] - hx;
  }
  float s = 0.f;
  for (int k = 0; k < m.dy; ++k) {
    float v = 0.f;
    for (int j = 0; j < m.dy; ++j) v += T[j][t] * m.SR[j * m.dy + k];
    s += v * v;
  }
  return 0.5f * s + m.nlz;
}

// phase 0: x = L0 eps + m0, lw = -lik(x, y) (0 without y); phase 1: x = A x_prev + LQ eps,
// lw = -lik(x, y); phase 2: lw = lik(x_prev, y). eps = normal(keys[i], (d,)).
__global__ __launch_bounds__(kLgThreads) void lg_hook_kernel(int phase, int d, int dy,
                                                             const float* __restrict__ params,
                                                             const float* __restrict__ xprev,
                                                             const float* __restrict__ y,
                                                             const uint32_t* __restrict__ keys, int64_t n,
                                                             float* __restrict__ xout, float* __restrict__ lw) {
  __shared__ float P[kLgMaxParams];
  __shared__ float E[kLgMaxDim][kLgThreads], XP[kLgMaxDim][kLgThreads], X[kLgMaxDim][kLgThreads];
  const int np = d + 3 * d * d + dy * d + dy * dy + 1;
  for (int k = threadIdx.x; k < np; k += kLgThreads) P[k] = params[k];
  __syncthreads();
  LgView m;
  m.d = d;
  m.dy = dy;
  m.m0 = P;
  m.L0 = m.m0 + d;
  m.A = m.L0 + d * d;
  m.LQ = m.A + d * d;
  m.H = m.LQ + d * d;
  m.SR = m.H + dy * d;
  m.nlz = m.SR[dy * dy];
  const int t = threadIdx.x;
  const uint32_t hd = ((uint32_t)d + 1u) >> 1;
  for (int64_t i = blockIdx.x * (int64_t)kLgThreads + t; i < n; i += (int64_t)gridDim.x * kLgThreads) {
    if (phase != 0)
      for (int q = 0; q < d; ++q) XP[q][t] = xprev[i * d + q];
    if (phase == 2) {
      lw[i] = lg_lik(m, XP, E, y, t);
      continue;
    }
    const Key ki{keys[2 * i], keys[2 * i + 1]};
    for (uint32_t q = 0; q < hd; ++q) {
      uint32_t w0, w1;
      bits_pair(ki, (uint32_t)d, q, w0, w1);
      E[q][t] = normal_from_bits(w0);
      if (q + hd < (uint32_t)d) E[q + hd][t] = normal_from_bits(w1);
    }
    float l = 0.f;
    if (phase == 0) {
      for (int r = 0; r < d; ++r) {
        float a = 0.f;
        for (int k = 0; k < d; ++k) a += m.L0[r * d + k] * E[k][t];
        X[r][t] = a + m.m0[r];
      }
      if (y) l = -lg_lik(m, X, E, y, t);
    } else {
      for (int r = 0; r < d; ++r) {
        float a = 0.f, b = 0.f;
        for (int k = 0; k < d; ++k) {
          a += m.A[r * d + k] * XP[k][t];
          b += m.LQ[r * d + k] * E[k][t];
        }
        X[r][t] = a + b;
      }
      l = -lg_lik(m, X, E, y, t);
    }
    for (int q = 0; q < d; ++q) xout[i * d + q] = X[q][t];
    lw[i] = l;
  }
}

}  // namespace

// Called by mocat_pf_hook (pf.hip) for model 0 at a (d, d_y) without a compiled kernel.
int launch_lg_generic_hook(int phase, int d, int dy, const float* params, const float* xprev, const float* y,
                           const uint32_t* keys, int64_t n, float* x, float* lw, hipStream_t st) {
  if (d < 1 || dy < 1 || d > kLgMaxDim || dy > kLgMaxDim) return MOCAT_ERR_ARG;
  const int64_t blocks = (n + kLgThreads - 1) / kLgThreads;
  hipLaunchKernelGGL(lg_hook_kernel, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(kLgThreads), 0, st, phase,
                     d, dy, params, xprev, y, keys, n, x, lw);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace mocat
