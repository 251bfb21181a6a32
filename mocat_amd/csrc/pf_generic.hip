// The linear-Gaussian model's per-particle functions at a runtime dimension, for the (d, d_y) outside the
// shapes pf.hip compiles, d and d_y up to kLgMaxDim:
//   model 0, the bootstrap filter: initial sample and weight, proposal and weight, likelihood potential
//     (mocat/src/ssm/linear_gaussian/linear_gaussian.py:46-51, 86-94, 118-128;
//     mocat/src/ssm/filtering.py:65-78, 142-170), noise normal(keys[i], (d,)) per particle;
//   model 2, the optimal proposal (mocat/src/ssm/nonlinear_gaussian.py:198-224, 264-281) from the
//     host's startup matrices, noise row i of normal(keys[0], (n, d)).
// One thread per particle, its vectors in LDS (rows indexed at run time), the parameters in LDS, and
// the formulas and summation order of pf.hip's Lgssm<D, DY> and LgssmOpt<D, DY>. mocat_pf_hook sends
// models 0 and 2 here when their (d, d_y) is not compiled; the host runs such a filter's time loop over
// these hooks and the library's resampling, gather and log-sum-exp kernels.
#include "common.hpp"

namespace mocat {

constexpr int kLgMaxDim = 16;

namespace {

constexpr int kLgThreads = 128;  // per-thread vectors: 4 x 16 x 128 floats = 32 KB of LDS per workgroup
// model 0: d + 3 d^2 + dy d + dy^2 + 1; model 2: d + 3 dy d + 3 d^2 + dy^2 + 1
constexpr int kLgMaxParams = kLgMaxDim + 7 * kLgMaxDim * kLgMaxDim + 1;

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

// Model 0: phase 0: x = L0 eps + m0, lw = -lik(x, y) (0 without y); phase 1: x = A x_prev + LQ eps,
// lw = -lik(x, y); phase 2: lw = lik(x_prev, y); eps = normal(keys[i], (d,)).
// Model 2 (params m0, H, K0, C0, A, K, S, WC, nlzw): phase 0: x = (m0 + K0 (y - H m0)) + eps @ C0,
// lw = 0; phase 1: mx = A x_prev, r = y - H mx, x = (mx + K r) + S eps, lw = -(0.5 |r @ WC|^2 + nlzw);
// eps = row i of normal(keys[0], (n, d)).
__global__ __launch_bounds__(kLgThreads) void lg_hook_kernel(int model, int phase, int d, int dy,
                                                             const float* __restrict__ params,
                                                             const float* __restrict__ xprev,
                                                             const float* __restrict__ y,
                                                             const uint32_t* __restrict__ keys, int64_t n,
                                                             float* __restrict__ xout, float* __restrict__ lw) {
  __shared__ float P[kLgMaxParams];
  __shared__ float E[kLgMaxDim][kLgThreads], XP[kLgMaxDim][kLgThreads], X[kLgMaxDim][kLgThreads];
  __shared__ float R[kLgMaxDim][kLgThreads];
  const int np = model == 2 ? d + 3 * dy * d + 3 * d * d + dy * dy + 1 : d + 3 * d * d + dy * d + dy * dy + 1;
  for (int k = threadIdx.x; k < np; k += kLgThreads) P[k] = params[k];
  __syncthreads();
  const int t = threadIdx.x;
  if (model == 2) {
    const float *m0 = P, *H = m0 + d, *K0 = H + dy * d, *C0 = K0 + d * dy, *A = C0 + d * d, *K = A + d * d,
                *S = K + d * dy, *WC = S + d * d;
    const float nlzw = WC[dy * dy];
    const Key k0{keys[0], keys[1]};
    const uint32_t total = (uint32_t)(n * d);
    for (int64_t i = blockIdx.x * (int64_t)kLgThreads + t; i < n; i += (int64_t)gridDim.x * kLgThreads) {
      for (int q = 0; q < d; ++q) E[q][t] = normal_from_bits(bits_at(k0, total, (uint32_t)(i * d + q)));
      if (phase == 0) {
        for (int j = 0; j < dy; ++j) {
          float hm = 0.f;
          for (int k = 0; k < d; ++k) hm += H[j * d + k] * m0[k];
          R[j][t] = y[j] - hm;
        }
        for (int r = 0; r < d; ++r) {
          float g = 0.f, c = 0.f;
          for (int j = 0; j < dy; ++j) g += K0[r * dy + j] * R[j][t];
          for (int k = 0; k < d; ++k) c += E[k][t] * C0[k * d + r];
          X[r][t] = (m0[r] + g) + c;
        }
        for (int q = 0; q < d; ++q) xout[i * d + q] = X[q][t];
        lw[i] = 0.f;
        continue;
      }
      for (int q = 0; q < d; ++q) XP[q][t] = xprev[i * d + q];
      for (int r = 0; r < d; ++r) {  // mx, in X (each row of X is written after its own row is read below)
        float a = 0.f;
        for (int k = 0; k < d; ++k) a += A[r * d + k] * XP[k][t];
        X[r][t] = a;
      }
      for (int j = 0; j < dy; ++j) {
        float hx = 0.f;
        for (int k = 0; k < d; ++k) hx += H[j * d + k] * X[k][t];
        R[j][t] = y[j] - hx;
      }
      for (int r = 0; r < d; ++r) {
        float g = 0.f, b = 0.f;
        for (int j = 0; j < dy; ++j) g += K[r * dy + j] * R[j][t];
        for (int k = 0; k < d; ++k) b += S[r * d + k] * E[k][t];
        X[r][t] = (X[r][t] + g) + b;
      }
      float s2 = 0.f;
      for (int k = 0; k < dy; ++k) {
        float v = 0.f;
        for (int j = 0; j < dy; ++j) v += R[j][t] * WC[j * dy + k];
        s2 += v * v;
      }
      for (int q = 0; q < d; ++q) xout[i * d + q] = X[q][t];
      lw[i] = -(0.5f * s2 + nlzw);
    }
    return;
  }
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

// The ensemble Kalman filter's member updates at a runtime (d, d_y) (mocat/src/ssm/nonlinear_gaussian.py:
// 297-350, EnsembleKalmanFilter on TimeHomogenousLinearGaussian's transition x -> A x), params as
// mocat_enkf_run's [m0, H, K0, C0, A, LQ, LR]; z = rows of normal(key, (n, k)) (k = d, or d_y for the
// observation perturbation), and the products in enkf.hip's order:
//   phase 0: x_i = (m0 + K0 (y - H m0)) + z_i @ C0
//   phase 1: mx_i = A x_i + LQ z_i (the forecast; xin = x, xout = mx)
//   phase 2: x_i = mx_i + K (y - (H mx_i + LR z_i)) with the gain K [d][d_y] (xin = mx)
__global__ __launch_bounds__(kLgThreads) void lg_enkf_kernel(int phase, int d, int dy, const float* __restrict__ params,
                                                             const float* __restrict__ xin, const uint32_t* __restrict__ key,
                                                             int64_t n, const float* __restrict__ y,
                                                             const float* __restrict__ gain, float* __restrict__ xout) {
  __shared__ float P[kLgMaxParams];
  __shared__ float Kg[kLgMaxDim * kLgMaxDim];
  __shared__ float Z[kLgMaxDim][kLgThreads], XI[kLgMaxDim][kLgThreads], R[kLgMaxDim][kLgThreads];
  const int np = d + 2 * dy * d + 3 * d * d + dy * dy;
  for (int k = threadIdx.x; k < np; k += kLgThreads) P[k] = params[k];
  if (phase == 2)
    for (int k = threadIdx.x; k < d * dy; k += kLgThreads) Kg[k] = gain[k];
  __syncthreads();
  const float *m0 = P, *H = m0 + d, *K0 = H + dy * d, *C0 = K0 + d * dy, *A = C0 + d * d, *LQ = A + d * d,
              *LR = LQ + d * d;
  const int t = threadIdx.x;
  const Key k{key[0], key[1]};
  const int kz = phase == 2 ? dy : d;  // noise columns
  const uint32_t total = (uint32_t)(n * kz);
  for (int64_t i = blockIdx.x * (int64_t)kLgThreads + t; i < n; i += (int64_t)gridDim.x * kLgThreads) {
    for (int q = 0; q < kz; ++q) Z[q][t] = normal_from_bits(bits_at(k, total, (uint32_t)(i * kz + q)));
    if (phase == 0) {
      for (int j = 0; j < dy; ++j) {
        float hm = 0.f;
        for (int q = 0; q < d; ++q) hm += H[j * d + q] * m0[q];
        R[j][t] = y[j] - hm;
      }
      for (int r = 0; r < d; ++r) {
        float g = 0.f, c = 0.f;
        for (int j = 0; j < dy; ++j) g += K0[r * dy + j] * R[j][t];
        for (int q = 0; q < d; ++q) c += Z[q][t] * C0[q * d + r];
        xout[i * d + r] = (m0[r] + g) + c;
      }
      continue;
    }
    for (int q = 0; q < d; ++q) XI[q][t] = xin[i * d + q];
    if (phase == 1) {
      for (int r = 0; r < d; ++r) {
        float u = 0.f, v = 0.f;
        for (int q = 0; q < d; ++q) {
          u += A[r * d + q] * XI[q][t];
          v += LQ[r * d + q] * Z[q][t];
        }
        xout[i * d + r] = u + v;
      }
      continue;
    }
    for (int j = 0; j < dy; ++j) {
      float hx = 0.f, e = 0.f;
      for (int q = 0; q < d; ++q) hx += H[j * d + q] * XI[q][t];
      for (int q = 0; q < dy; ++q) e += LR[j * dy + q] * Z[q][t];
      R[j][t] = y[j] - (hx + e);
    }
    for (int r = 0; r < d; ++r) {
      float g = 0.f;
      for (int j = 0; j < dy; ++j) g += Kg[r * dy + j] * R[j][t];
      xout[i * d + r] = XI[r][t] + g;
    }
  }
}

}  // namespace

// Called by mocat_pf_hook (pf.hip) for models 0 and 2 at a (d, d_y) without a compiled kernel (its
// argument checks done there: model 2 takes phases 0 and 1 with y, n * d < 2^32).
int launch_lg_generic_hook(int model, int phase, int d, int dy, const float* params, const float* xprev,
                           const float* y, const uint32_t* keys, int64_t n, float* x, float* lw, hipStream_t st) {
  if ((model != 0 && model != 2) || d < 1 || dy < 1 || d > kLgMaxDim || dy > kLgMaxDim) return MOCAT_ERR_ARG;
  if (model == 2 && (phase == 2 || !y)) return MOCAT_ERR_ARG;
  const int64_t blocks = (n + kLgThreads - 1) / kLgThreads;
  hipLaunchKernelGGL(lg_hook_kernel, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(kLgThreads), 0, st, model,
                     phase, d, dy, params, xprev, y, keys, n, x, lw);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace mocat

extern "C" {

int mocat_enkf_hook(int phase, int d, int dy, const float* params, const float* x_in, const uint32_t* key, int64_t n,
                    const float* y, const float* gain, float* x_out, hipStream_t stream) {
  using namespace mocat;
  if (phase < 0 || phase > 2 || d < 1 || dy < 1 || d > kLgMaxDim || dy > kLgMaxDim || !params || !key || !x_out ||
      n <= 0 || (phase != 1 && !y) || (phase != 0 && !x_in) || (phase == 2 && !gain))
    return MOCAT_ERR_ARG;
  if (n * (int64_t)(d > dy ? d : dy) >= ((int64_t)1 << 32)) return MOCAT_ERR_SIZE;
  const int64_t blocks = (n + kLgThreads - 1) / kLgThreads;
  hipLaunchKernelGGL(lg_enkf_kernel, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(kLgThreads), 0, stream,
                     phase, d, dy, params, x_in, key, n, y, gain, x_out);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
