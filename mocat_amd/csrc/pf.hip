// Bootstrap particle filters and resampling.
//
// pf_kernel: one persistent workgroup per filter runs the whole time loop of mocat's
// run_particle_filter_for_marginals (mocat/src/ssm/filtering.py:255-324) for that filter:
//   init   (key, sub) = split(key); x_i ~ initial_sample(split(sub, n)[i]); lw_i = -lik(x_i, y_0)
//          (filtering.py:173-193, 264-270)
//   step t int_key = split(key, T-1)[t-1]; resample iff ESS_{t-1} < threshold * n
//          (filtering.py:284-292); x_i = transition_sample(x_{a_i}, split(int_key, n)[i]);
//          lw_i = lw_res_i - lik(x_i, y_t) (filtering.py:294-303, BootstrapFilter :142-170).
// The log-likelihood estimate (not returned by mocat) accumulates in FP64:
//   log Z = sum_t [lse(lw_t) - lse(lw_res_{t-1})], lse(lw_res) = log n after a resample,
//   t = 0 term lse(lw_0) - log n.
// Filters of a batch may have different n and T (the host orders them by n*T, largest first).
//
// Resampling given weights w (any non-negative f32) is exact and order independent: weights are
// scaled by a power of two and truncated to integers q_i (no rounding: the scale is exact), their
// prefix sums C_i are exact 64-bit integers, and a target T_j selects the first i with C_i > T_j:
//   systematic   T_j = (j + u) * (W / n)     u = uniform(key), W = C_{n-1} (FP64, IEEE rounding)
//   multinomial  T_j = u_j * W               u = uniform(key, (n,))
// The reference's resampler, random.categorical = Gumbel-max over an (n, n) Gumbel matrix
// (filtering.py:196-199), is also provided (O(n^2)).
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "common.hpp"

using namespace mocat;

namespace mocat {
// pf_generic.hip: the linear-Gaussian hooks (models 0 and 2) at a runtime (d, d_y), both up to 16
int launch_lg_generic_hook(int model, int phase, int d, int dy, const float* params, const float* xprev,
                           const float* y, const uint32_t* keys, int64_t n, float* x, float* lw, hipStream_t st);
}  // namespace mocat

extern "C" {
typedef struct {
  int64_t n;
  int64_t T;
  int64_t y_offset;
  int64_t state_offset;
  int64_t out_offset;
  int64_t t_offset;
  uint32_t key[2];
  int32_t param_offset;
  int32_t flags;
} mocat_pf_filter;
}

namespace {

constexpr int NT = 512;
constexpr int NW = NT / kWave;
constexpr int32_t kPfLarge = 1;  // mocat_pf_filter.flags: run by mocat_pf_run_large, skipped by mocat_pf_run
// mocat_pf_filter.flags: resume from a given sample (run_particle_filter_for_marginals with initial_sample,
// mocat/src/ssm/filtering.py:264-278): the particles and log-weights of t = 0 are read from the state
// workspaces x_ws0 / lw_ws0 instead of drawn, the step keys are split(key, T - 1) (no (key, sub) split)
// and the log-likelihood starts at 0 (y_0 is not read).
constexpr int32_t kPfResume = 2;
// mocat_pf_filter.flags for mocat's one-step functions (filtering.py:173-253): step t's key is the
// filter key itself (propagate_particle_filter: T = 2), and the initial particles come from
// split(key, n) without the (key, sub) split (initiate_particles).
constexpr int32_t kPfStepKey = 4;
constexpr int32_t kPfInitKey = 8;
constexpr double kLog2PiD = 1.83787706640934548356;

// A compiler-only memory fence (no instruction). Placed at the top of a per-particle loop, it stops
// the compiler from hoisting the model's LDS-resident matrices into registers for the whole loop
// (at D = 8 that took more than the 256 VGPRs a 512-thread workgroup leaves and spilled to scratch);
// each iteration re-reads them from LDS instead. At D >= 8 the models also fence between the rows
// of each matrix product, so that no more than a few rows' loads are in flight at once.
MOCAT_DEV void reload_model_each_particle() { asm volatile("" ::: "memory"); }
template <int D>
MOCAT_DEV void row_fence() {
  if constexpr (D >= 8) asm volatile("" ::: "memory");
}

// ---------------------------------------------------------------- models
// Linear Gaussian (mocat/src/ssm/linear_gaussian/linear_gaussian.py:46-51, 86-94, 118-128):
// params: m0[D], L0[D*D], A[D*D], LQ[D*D], H[DY*D], SR[DY*DY], nlz
//   x0 = L0 @ eps + m0 ; x' = A @ x + LQ @ eps ; lik = 0.5 |(y - H x) @ SR|^2 + nlz
template <int D, int DY>
struct Lgssm {
  static constexpr bool kOptimal = false;
  static constexpr bool kTimeScaled = false;  // time homogeneous: the observation times are not read
  static constexpr int NPARAM = D + 3 * D * D + DY * D + DY * DY + 1;
  float m0[D], L0[D * D], A[D * D], LQ[D * D], H[DY * D], SR[DY * DY], nlz;
  MOCAT_DEV void load(const float* p) {
    #pragma unroll
    for (int i = 0; i < D; ++i) m0[i] = p[i];
    p += D;
    #pragma unroll
    for (int i = 0; i < D * D; ++i) L0[i] = p[i];
    p += D * D;
    #pragma unroll
    for (int i = 0; i < D * D; ++i) A[i] = p[i];
    p += D * D;
    #pragma unroll
    for (int i = 0; i < D * D; ++i) LQ[i] = p[i];
    p += D * D;
    #pragma unroll
    for (int i = 0; i < DY * D; ++i) H[i] = p[i];
    p += DY * D;
    #pragma unroll
    for (int i = 0; i < DY * DY; ++i) SR[i] = p[i];
    p += DY * DY;
    nlz = p[0];
  }
  MOCAT_DEV void init(const float* eps, float* x) const {
    #pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f;
      #pragma unroll
      for (int k = 0; k < D; ++k) a += L0[i * D + k] * eps[k];
      x[i] = a + m0[i];
      row_fence<D>();
    }
  }
  MOCAT_DEV void transition(const float* xp, const float* eps, float* x) const {
    #pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f, b = 0.f;
      #pragma unroll
      for (int k = 0; k < D; ++k) {
        a += A[i * D + k] * xp[k];
        b += LQ[i * D + k] * eps[k];
      }
      x[i] = a + b;
      row_fence<D>();
    }
  }
  MOCAT_DEV float lik(const float* x, const float* y) const {
    float diff[DY];
    #pragma unroll
    for (int j = 0; j < DY; ++j) {
      float hx = 0.f;
      #pragma unroll
      for (int k = 0; k < D; ++k) hx += H[j * D + k] * x[k];
      diff[j] = y[j] - hx;
      row_fence<D>();
    }
    float s = 0.f;
    #pragma unroll
    for (int k = 0; k < DY; ++k) {
      float v = 0.f;
      #pragma unroll
      for (int j = 0; j < DY; ++j) v += diff[j] * SR[j * DY + k];
      s += v * v;
      row_fence<D>();
    }
    return 0.5f * s + nlz;
  }
};

// Stochastic volatility, 1-D (mocat/src/ssm/scenarios/stochastic_volatility.py:61-108), unit steps:
// params: m0, L0, phi, sigma, inv_beta (likelihood precision sqrt 1/beta)
//   x0 = eps * L0 + m0 ; x' = phi x + sigma eps ;
//   lik = gaussian_potential(y, 0, sqrt_prec = inv_beta / exp(x/2))
//       = 0.5 (y s)^2 + 0.5 log 2pi - log s,  s = inv_beta * exp(-x/2)
struct Sv {
  static constexpr bool kOptimal = false;
  // the transition noise scales with sqrt(t_new - t_previous) (stochastic_volatility.py:86-93)
  static constexpr bool kTimeScaled = true;
  static constexpr int NPARAM = 5;
  float m0, L0, phi, sigma, inv_beta;
  MOCAT_DEV void load(const float* p) {
    m0 = p[0];
    L0 = p[1];
    phi = p[2];
    sigma = p[3];
    inv_beta = p[4];
  }
  MOCAT_DEV void init(const float* eps, float* x) const { x[0] = eps[0] * L0 + m0; }
  MOCAT_DEV void transition(const float* xp, const float* eps, float* x) const { x[0] = xp[0] * phi + eps[0] * sigma; }
  // with the step's noise scale s = sqrt(t_new - t_previous): x' = phi x + (sigma eps) s, the product in
  // mocat's order (normal @ cov_sqrt.T * sqrt(dt)); s = 1 gives transition() bit for bit
  MOCAT_DEV void transition_s(const float* xp, const float* eps, float* x, float s) const {
    x[0] = xp[0] * phi + (eps[0] * sigma) * s;
  }
  // one expression for both paths, so that computing the gradient leaves the log-weights bit-identical
  MOCAT_DEV float lik(const float* x, const float* y) const {
    float dldx, dldb;
    return lik_grad(x, y, dldx, dldb);
  }
  // d lik / dx = 0.5 - 0.5 (y s)^2 ; d lik / dbeta = (1 - (y s)^2) * inv_beta
  MOCAT_DEV float lik_grad(const float* x, const float* y, float& dldx, float& dldb) const {
    const float s = inv_beta / __expf(0.5f * x[0]);
    const float ys = y[0] * s;
    const float s2 = ys * ys;
    dldx = 0.5f - 0.5f * s2;
    dldb = (1.f - s2) * inv_beta;
    return 0.5f * s2 + 0.5f * (float)kLog2PiD - __logf(s);
  }
};

// Multivariate stochastic volatility, dim D = dim_obs (mocat/src/ssm/scenarios/stochastic_volatility.py:15-108):
// params: m0[D], L0[D*D], A[D*D], LQ[D*D], ib[D] (L0, LQ Cholesky factors of the initial and transition
// covariances, ib_k = 1 / sqrt(likelihood_covariance_diag_k))
//   x0 = L0 eps + m0 ; x' = A x + (LQ eps) s with s = sqrt(t_new - t_previous) ;
//   lik = gaussian_potential(y, 0, sqrt_prec = ib / exp(x / 2)) = sum_k 0.5 (y_k s_k)^2 + 0.5 log 2pi - log s_k
template <int D>
struct SvD {
  static constexpr bool kOptimal = false;
  static constexpr bool kTimeScaled = true;
  static constexpr int NPARAM = D + 3 * D * D + D;
  float m0[D], L0[D * D], A[D * D], LQ[D * D], ib[D];
  MOCAT_DEV void load(const float* p) {
#pragma unroll
    for (int i = 0; i < D; ++i) m0[i] = p[i];
    p += D;
#pragma unroll
    for (int i = 0; i < D * D; ++i) L0[i] = p[i];
    p += D * D;
#pragma unroll
    for (int i = 0; i < D * D; ++i) A[i] = p[i];
    p += D * D;
#pragma unroll
    for (int i = 0; i < D * D; ++i) LQ[i] = p[i];
    p += D * D;
#pragma unroll
    for (int i = 0; i < D; ++i) ib[i] = p[i];
  }
  MOCAT_DEV void init(const float* eps, float* x) const {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) a += L0[i * D + k] * eps[k];
      x[i] = a + m0[i];
    }
  }
  MOCAT_DEV void transition_s(const float* xp, const float* eps, float* x, float s) const {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f, b = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        a += A[i * D + k] * xp[k];
        b += LQ[i * D + k] * eps[k];
      }
      x[i] = a + b * s;
    }
  }
  MOCAT_DEV void transition(const float* xp, const float* eps, float* x) const { transition_s(xp, eps, x, 1.f); }
  MOCAT_DEV float lik(const float* x, const float* y) const {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const float sk = ib[k] / __expf(0.5f * x[k]);
      const float ys = y[k] * sk;
      acc += 0.5f * (ys * ys) + 0.5f * (float)kLog2PiD - __logf(sk);
    }
    return acc;
  }
};

// Optimal proposal for the linear-Gaussian model: OptimalNonLinearGaussianParticleFilter
// (mocat/src/ssm/nonlinear_gaussian.py:134-281) with transition_function(x) = A x
// (linear_gaussian.py:112-116). Matrices factored on the host as its startup (:154-184):
// params: m0[D], H[DY*D], K0[D*DY], C0[D*D], A[D*D], K[D*DY], S[D*D], WC[DY*DY], nlzw
//   x0 = m0 + K0 (y0 - H m0) + eps @ C0, lw0 = 0                                   (:198-224)
//   mx = A x ; r = y - H mx ; x' = mx + K r + S eps ; lw += -(0.5 |r @ WC|^2 + nlzw)  (:265-281)
// All particles' noise comes from one key per step: normal(split(key, n)[0], (n, D)).
template <int D, int DY>
struct LgssmOpt {
  static constexpr bool kOptimal = true;
  static constexpr bool kTimeScaled = false;
  static constexpr int NPARAM = D + 3 * DY * D + 3 * D * D + DY * DY + 1;
  float m0[D], H[DY * D], K0[D * DY], C0[D * D], A[D * D], K[D * DY], S[D * D], WC[DY * DY], nlzw;
  MOCAT_DEV void load(const float* p) {
    auto take = [&](float* dst, int cnt) {
      for (int i = 0; i < cnt; ++i) dst[i] = p[i];
      p += cnt;
    };
    take(m0, D);
    take(H, DY * D);
    take(K0, D * DY);
    take(C0, D * D);
    take(A, D * D);
    take(K, D * DY);
    take(S, D * D);
    take(WC, DY * DY);
    nlzw = p[0];
  }
  MOCAT_DEV void init_opt(const float* eps, const float* y0, float* x) const {
    float r[DY];
#pragma unroll
    for (int j = 0; j < DY; ++j) {
      float hm = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) hm += H[j * D + k] * m0[k];
      r[j] = y0[j] - hm;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float g = 0.f, c = 0.f;
#pragma unroll
      for (int j = 0; j < DY; ++j) g += K0[i * DY + j] * r[j];
#pragma unroll
      for (int k = 0; k < D; ++k) c += eps[k] * C0[k * D + i];
      x[i] = (m0[i] + g) + c;
    }
  }
  // returns the weight potential; x receives the proposed state
  MOCAT_DEV float step_opt(const float* xp, const float* eps, const float* y, float* x) const {
    float mx[D], r[DY];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) a += A[i * D + k] * xp[k];
      mx[i] = a;
    }
#pragma unroll
    for (int j = 0; j < DY; ++j) {
      float hx = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) hx += H[j * D + k] * mx[k];
      r[j] = y[j] - hx;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float g = 0.f, b = 0.f;
#pragma unroll
      for (int j = 0; j < DY; ++j) g += K[i * DY + j] * r[j];
#pragma unroll
      for (int k = 0; k < D; ++k) b += S[i * D + k] * eps[k];
      x[i] = (mx[i] + g) + b;
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < DY; ++k) {
      float v = 0.f;
#pragma unroll
      for (int j = 0; j < DY; ++j) v += r[j] * WC[j * DY + k];
      s += v * v;
    }
    return 0.5f * s + nlzw;
  }
};

// Row `row` of normal(k0, (n, D)) (S = n * D elements in jax's halves layout).
template <int D>
MOCAT_DEV void shared_eps(Key k0, uint32_t S, uint32_t row, float* eps) {
#pragma unroll
  for (int q = 0; q < D; ++q) eps[q] = normal_from_bits(bits_at(k0, S, row * (uint32_t)D + (uint32_t)q));
}

// ---------------------------------------------------------------- block helpers
struct LseAcc {
  float m;
  double s1, s2;
};

struct BlockScratch {
  float fm[NW];
  double d1[NW], d2[NW];
  uint64_t u64[NT];
  double g0[NW], g1[NW], g2[NW];
  float bcast_f[4];
  double bcast_d[4];
  uint32_t bcast_u[4];
};

// Reduce per-thread LseAcc across the block; returns (lse, lse2) to all threads.
// Log-sum-exp of the log-weights (and, with GRAD, the weighted means of the tangents) over the
// thread's particles i = tid, tid + NT, ... < n, whose log-weights (and tangents, 5 floats per
// particle, the last three taken) the thread has just written: the workgroup maximum of the values (m_local: the thread's maximum,
// NaN excluded; bad: a NaN was seen), then one FP64 exponential per particle and plain FP64 sums,
// lanes by butterflies and waves in order.
template <bool GRAD>
MOCAT_DEV void block_lse2(const float* lw, const float* tan, int64_t n, float m_local, int bad, BlockScratch& s,
                          double& lse, double& lse2, double* gout) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  float m = wave_max(m_local);
  bad = __any(bad) ? 1 : 0;
  __syncthreads();
  if (lane == 0) {
    s.fm[w] = m;
    s.u64[w] = (uint64_t)bad;
  }
  __syncthreads();
  float M = s.fm[0];
  bool anybad = s.u64[0] != 0;
  for (int k = 1; k < NW; ++k) {
    M = fmaxf(M, s.fm[k]);
    anybad = anybad || s.u64[k] != 0;
  }
  double s1 = 0.0, s2 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (M != -INFINITY && !anybad) {
    for (int64_t i = threadIdx.x; i < n; i += NT) {
      const float v = lw[i];
      if (!(v > -INFINITY)) continue;
      const double e = exp((double)v - (double)M);
      s1 += e;
      s2 += e * e;
      if constexpr (GRAD) {
        g0 += e * tan[i * 5 + 2];
        g1 += e * tan[i * 5 + 3];
        g2 += e * tan[i * 5 + 4];
      }
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if constexpr (GRAD) {
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
  }
  __syncthreads();
  if (lane == 0) {
    s.d1[w] = s1;
    s.d2[w] = s2;
    if constexpr (GRAD) {
      s.g0[w] = g0;
      s.g1[w] = g1;
      s.g2[w] = g2;
    }
  }
  __syncthreads();
  double r1 = s.d1[0], r2 = s.d2[0], q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if constexpr (GRAD) {
    q0 = s.g0[0];
    q1 = s.g1[0];
    q2 = s.g2[0];
  }
  for (int k = 1; k < NW; ++k) {
    r1 += s.d1[k];
    r2 += s.d2[k];
    if constexpr (GRAD) {
      q0 += s.g0[k];
      q1 += s.g1[k];
      q2 += s.g2[k];
    }
  }
  if (anybad) {
    lse = NAN;
    lse2 = NAN;
  } else if (M == -INFINITY) {
    lse = -INFINITY;
    lse2 = -INFINITY;
  } else {
    lse = (double)M + log(r1);
    lse2 = 2.0 * (double)M + log(r2);
  }
  if constexpr (GRAD) {
    gout[0] = q0 / r1;
    gout[1] = q1 / r1;
    gout[2] = q2 / r1;
  }
}

// Exclusive block scan of one uint64 per thread; returns the block total to all threads.
MOCAT_DEV uint64_t block_exclusive_scan(uint64_t v, BlockScratch& s, uint64_t& total) {
  __syncthreads();
  s.u64[threadIdx.x] = v;
  __syncthreads();
  // Hillis-Steele in LDS (log2 NT rounds)
  for (int o = 1; o < NT; o <<= 1) {
    const uint64_t add = threadIdx.x >= o ? s.u64[threadIdx.x - o] : 0ull;
    __syncthreads();
    s.u64[threadIdx.x] += add;
    __syncthreads();
  }
  total = s.u64[NT - 1];
  const uint64_t incl = s.u64[threadIdx.x];
  __syncthreads();
  return incl - v;
}

// Quantisation shift: weights scaled by 2^shift with the largest < 2^qb, qb leaves headroom for n terms.
MOCAT_DEV int quant_shift(float wmax, int64_t n) {
  int e;
  frexpf(wmax, &e);  // wmax = f * 2^e, f in [0.5, 1)
  int bits_n = 0;
  while ((1ll << bits_n) <= n) ++bits_n;
  const int qb = 52 - bits_n;
  return qb - e;
}

MOCAT_DEV uint64_t quantise(float w, int shift) {
  if (!(w > 0.f)) return 0ull;
  return (uint64_t)ldexp((double)w, shift);  // exact scaling, truncation toward zero
}

// First index i in [0, n) with C[i] > t (C inclusive prefix sums as FP64 integers); n - 1 if none.
MOCAT_DEV int64_t search_csum(const uint64_t* __restrict__ C, int64_t n, double t) {
  int64_t lo = 0, hi = n;  // invariant: answer in [lo, hi]
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((double)C[mid] > t)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo < n ? lo : n - 1;
}

// Inclusive prefix sums of the quantised weights w_i = weight(i) into csum (n uint64), by one
// workgroup (all threads participate); returns the total.
template <class WeightFn>
MOCAT_DEV uint64_t quantised_csum_block(WeightFn weight, int64_t n, uint64_t* __restrict__ csum, BlockScratch& s) {
  const int t = threadIdx.x;
  // max weight
  float wmax = 0.f;
  for (int64_t i = t; i < n; i += NT) wmax = fmaxf(wmax, weight(i));
  wmax = block_max(wmax, s.fm);
  const int shift = quant_shift(wmax, n);
  // contiguous chunk per thread: local sums, block scan, then prefix sums written out
  const int64_t chunk = (n + NT - 1) / NT;
  const int64_t b = (int64_t)t * chunk, e = b + chunk < n ? b + chunk : n;
  uint64_t local = 0;
  for (int64_t i = b; i < e; ++i) local += quantise(weight(i), shift);
  uint64_t total;
  uint64_t run = block_exclusive_scan(local, s, total);
  for (int64_t i = b; i < e; ++i) {
    run += quantise(weight(i), shift);
    csum[i] = run;
  }
  return total;
}

// The one uniform of systematic resampling.
MOCAT_DEV double systematic_uniform(Key key) { return (double)uniform_from_bits(bits_at(key, 1, 0), 0.f, 1.f); }

// Ancestor of output j from the exact prefix sums (total W): systematic (method 0, u_sys =
// systematic_uniform(key)) or multinomial (method 1, uniform j of n from key).
MOCAT_DEV int32_t ancestor_at(const uint64_t* __restrict__ csum, int64_t n, double W, Key key, int method,
                              double u_sys, int64_t j) {
  if (method == 0) return (int32_t)search_csum(csum, n, ((double)j + u_sys) * (W / (double)n));
  const double u = (double)uniform_from_bits(bits_at(key, (uint32_t)n, (uint32_t)j), 0.f, 1.f);
  return (int32_t)search_csum(csum, n, u * W);
}

// Ancestor indices for weights w_i = weight(i) (all threads participate).
// method 0 systematic, 1 multinomial (uniforms from key). csum: n uint64 of workspace.
template <class WeightFn>
MOCAT_DEV void resample_block(WeightFn weight, int64_t n, Key key, int method, uint64_t* __restrict__ csum,
                              int32_t* __restrict__ anc, BlockScratch& s) {
  const uint64_t total = quantised_csum_block(weight, n, csum, s);
  __threadfence_block();
  __syncthreads();
  const double W = (double)total, u_sys = method == 0 ? systematic_uniform(key) : 0.0;
  for (int64_t j = threadIdx.x; j < n; j += NT) anc[j] = ancestor_at(csum, n, W, key, method, u_sys, j);
  __syncthreads();
}

// random.categorical(key, lw, shape=(n_out,)): argmax_i lw_i + gumbel(key, (n_out, n))[j, i]
// (first maximum wins), gumbel = -log(-log(uniform(minval=tiny, maxval=1))).
MOCAT_DEV int32_t categorical_row(const float* __restrict__ lw, int64_t n, Key key, uint32_t S, int64_t j) {
  float best = -INFINITY;
  int32_t bi = 0;
  for (int64_t i = 0; i < n; ++i) {
    const float u = uniform_from_bits(bits_at(key, S, (uint32_t)(j * n + i)), 1.17549435e-38f, 1.f);
    const float g = -logf(-logf(u)) + lw[i];
    if (g > best || (i == 0)) {
      if (i == 0 || g > best) {
        best = g;
        bi = (int32_t)i;
      }
    }
  }
  return bi;
}

// ---------------------------------------------------------------- the persistent filter kernel
struct PfArgs {
  const mocat_pf_filter* filters;
  int n_filters;
  const float* params;
  const float* y;
  const float* step_scale;  // [rows of y] noise scale of the step into row t (time-scaled models), or NULL
  float* x0;       // workspace [sum n][D]
  float* x1;       // workspace [sum n][D]
  float* lw0;      // workspace [sum n]
  float* lw1;      // workspace [sum n]
  uint64_t* csum;  // workspace [sum n]
  int32_t* anc;    // workspace [sum n]
  int resampling;  // 0 systematic, 1 multinomial, 2 categorical
  float ess_threshold;
  float* out_value;   // [T][n][D] at out_offset*D, or NULL
  float* out_lw;      // [T][n] at out_offset, or NULL
  int32_t* out_anc;   // [T][n] at out_offset (row 0 = identity), or NULL
  float* out_ess;     // [T] at t_offset, or NULL
  double* out_loglik_t;  // [T] at t_offset: cumulative log-likelihood after step t, or NULL
  double* out_loglik;    // [n_filters]
  float* tan0;           // [sum n][5] tangents (dx/dphi, dx/dsigma, dlw/dphi, dlw/dsigma, dlw/dbeta), ping
  float* tan1;           // pong
  double* out_grad;      // [n_filters][3] d log Z / d(phi, sigma, beta) (stochastic volatility only)
};

// GRAD (stochastic volatility only): forward tangents of the fixed-resampling-index estimator,
// d log Z / d(phi, sigma, beta): per particle dx/dphi, dx/dsigma (gathered by ancestor like x) and the
// log-weight tangents; d log Z = sum_t [avg_w(dlw_t) - avg_w(dlw_res)] in FP64.
template <class Model, int D, int DY, bool GRAD>
__global__ __launch_bounds__(NT) void pf_kernel(PfArgs a) {
  __shared__ BlockScratch s;
  __shared__ Model model;
  const int tid = threadIdx.x;
  for (int f = blockIdx.x; f < a.n_filters; f += gridDim.x) {
    const mocat_pf_filter F = a.filters[f];
    if (F.flags & kPfLarge) continue;
    const int64_t n = F.n, T = F.T;
    __syncthreads();
    load_shared_params(model, a.params + F.param_offset);
    __syncthreads();
    float* xa = a.x0 + F.state_offset * D;
    float* xb = a.x1 + F.state_offset * D;
    float* lwa = a.lw0 + F.state_offset;
    float* lwb = a.lw1 + F.state_offset;
    uint64_t* csum = a.csum + F.state_offset;
    int32_t* anc = a.anc + F.state_offset;
    const float* yf = a.y + F.y_offset;
    const Key key{F.key[0], F.key[1]};
    const bool resume = (F.flags & kPfResume) != 0;
    const Key rkey = resume ? key : split_at(key, 2, 0);
    const Key sub = (F.flags & kPfInitKey) ? key : split_at(key, 2, 1);
    constexpr uint32_t hD = (D + 1) / 2;
    float* ta = GRAD ? a.tan0 + F.state_offset * 5 : nullptr;
    float* tb = GRAD ? a.tan1 + F.state_offset * 5 : nullptr;
    double gtot[3] = {0.0, 0.0, 0.0}, gprev[3] = {0.0, 0.0, 0.0};

    // ---- t = 0: initiate_particles
    float mloc = -INFINITY;
    int bad0 = 0;
    const Key sub0 = Model::kOptimal ? split_at(sub, (uint32_t)n, 0u) : Key{0u, 0u};
    for (int64_t i = tid; i < n; i += NT) {
      reload_model_each_particle();
      float eps[D], x[D];
      float l;
      if (resume) {
        #pragma unroll
        for (int q = 0; q < D; ++q) x[q] = xa[i * D + q];
        l = lwa[i];
      } else if constexpr (Model::kOptimal) {
        shared_eps<D>(sub0, (uint32_t)(n * D), (uint32_t)i, eps);
        model.init_opt(eps, yf, x);
        l = 0.f;
      } else {
      const Key ki = split_at(sub, (uint32_t)n, (uint32_t)i);
      #pragma unroll
      for (uint32_t q = 0; q < hD; ++q) {
        uint32_t w0, w1;
        bits_pair(ki, D, q, w0, w1);
        eps[q] = normal_from_bits(w0);
        if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
      }
      model.init(eps, x);
      if constexpr (GRAD) {
        float dldx, dldb;
        l = -model.lik_grad(x, yf, dldx, dldb);
        // the initial state does not depend on (phi, sigma)
        ta[i * 5 + 0] = 0.f;
        ta[i * 5 + 1] = 0.f;
        ta[i * 5 + 2] = 0.f;
        ta[i * 5 + 3] = 0.f;
        ta[i * 5 + 4] = -dldb;
      } else {
        l = -model.lik(x, yf);
      }
      }
      #pragma unroll
      for (int q = 0; q < D; ++q) xa[i * D + q] = x[q];
      lwa[i] = l;
      if (isnan(l))
        bad0 = 1;
      else
        mloc = fmaxf(mloc, l);
      if (a.out_value && F.out_offset >= 0)
        #pragma unroll
        for (int q = 0; q < D; ++q) a.out_value[(F.out_offset + i) * D + q] = x[q];
      if (a.out_lw && F.out_offset >= 0) a.out_lw[F.out_offset + i] = l;
      if (a.out_anc && F.out_offset >= 0) a.out_anc[F.out_offset + i] = (int32_t)i;
    }
    double lse, lse2;
    block_lse2<GRAD>(lwa, ta, n, mloc, bad0, s, lse, lse2, gprev);
    if constexpr (GRAD)
      for (int k = 0; k < 3; ++k) gtot[k] = gprev[k];
    const double logn = log((double)n);
    double loglik = resume ? 0.0 : lse - logn;
    double ess = exp(2.0 * lse - lse2);
    if (tid == 0) {
      if (a.out_ess) a.out_ess[F.t_offset] = (float)ess;
      if (a.out_loglik_t) a.out_loglik_t[F.t_offset] = loglik;
    }

    // ---- t = 1 .. T-1
    for (int64_t t = 1; t < T; ++t) {
      const Key ik = (F.flags & kPfStepKey) ? key : split_at(rkey, (uint32_t)(T - 1), (uint32_t)(t - 1));
      const bool res = ess < (double)a.ess_threshold * (double)n;
      const float* ycur = yf + t * DY;
      const double lse_res = res ? logn : lse;
      if (res) {
        if (a.resampling == 2) {
          const uint32_t S = (uint32_t)(n * n);
          for (int64_t j = tid; j < n; j += NT) anc[j] = categorical_row(lwa, n, ik, S, j);
          __syncthreads();
        } else {
          const float m = (float)(lse);  // any finite shift works: weights are rescaled exactly
          auto wfn = [&](int64_t i) { return __expf(lwa[i] - m); };
          resample_block(wfn, n, ik, a.resampling, csum, anc, s);
        }
      }
      mloc = -INFINITY;
      bad0 = 0;
      const Key ik0 = Model::kOptimal ? split_at(ik, (uint32_t)n, 0u) : Key{0u, 0u};
      float sc = 1.f;
      if constexpr (Model::kTimeScaled)
        if (a.step_scale) sc = a.step_scale[F.y_offset / DY + t];
      for (int64_t i = tid; i < n; i += NT) {
        reload_model_each_particle();
        const int64_t src = res ? (int64_t)anc[i] : i;
        float eps[D], xp[D], x[D];
        if constexpr (Model::kOptimal) {
          shared_eps<D>(ik0, (uint32_t)(n * D), (uint32_t)i, eps);
        } else {
          const Key ki = split_at(ik, (uint32_t)n, (uint32_t)i);
          #pragma unroll
          for (uint32_t q = 0; q < hD; ++q) {
            uint32_t w0, w1;
            bits_pair(ki, D, q, w0, w1);
            eps[q] = normal_from_bits(w0);
            if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
          }
        }
        #pragma unroll
        for (int q = 0; q < D; ++q) xp[q] = xa[src * D + q];
        float l;
        if constexpr (Model::kOptimal) {
          l = (res ? 0.f : lwa[i]) - model.step_opt(xp, eps, ycur, x);
        } else if constexpr (GRAD) {
          model.transition_s(xp, eps, x, sc);
          float dldx, dldb;
          l = (res ? 0.f : lwa[i]) - model.lik_grad(x, ycur, dldx, dldb);
          const float dxp = xp[0] + model.phi * ta[src * 5 + 0];
          const float dxs = eps[0] * sc + model.phi * ta[src * 5 + 1];
          const float d0 = (res ? 0.f : ta[i * 5 + 2]) - dldx * dxp;
          const float d1 = (res ? 0.f : ta[i * 5 + 3]) - dldx * dxs;
          const float d2 = (res ? 0.f : ta[i * 5 + 4]) - dldb;
          tb[i * 5 + 0] = dxp;
          tb[i * 5 + 1] = dxs;
          tb[i * 5 + 2] = d0;
          tb[i * 5 + 3] = d1;
          tb[i * 5 + 4] = d2;

        } else {
          if constexpr (Model::kTimeScaled)
            model.transition_s(xp, eps, x, sc);
          else
            model.transition(xp, eps, x);
          l = (res ? 0.f : lwa[i]) - model.lik(x, ycur);
        }
        #pragma unroll
        for (int q = 0; q < D; ++q) xb[i * D + q] = x[q];
        lwb[i] = l;
        if (isnan(l))
          bad0 = 1;
        else
          mloc = fmaxf(mloc, l);
        if (F.out_offset >= 0) {
          const int64_t o = F.out_offset + t * n + i;
          if (a.out_value)
            #pragma unroll
            for (int q = 0; q < D; ++q) a.out_value[o * D + q] = x[q];
          if (a.out_lw) a.out_lw[o] = l;
          if (a.out_anc) a.out_anc[o] = (int32_t)src;
        }
      }
      double gnow[3];
      block_lse2<GRAD>(lwb, tb, n, mloc, bad0, s, lse, lse2, gnow);
      loglik += lse - lse_res;
      if constexpr (GRAD) {
        for (int k = 0; k < 3; ++k) {
          gtot[k] += gnow[k] - (res ? 0.0 : gprev[k]);
          gprev[k] = gnow[k];
        }
        float* tt = ta;
        ta = tb;
        tb = tt;
      }
      ess = exp(2.0 * lse - lse2);
      if (tid == 0) {
        if (a.out_ess) a.out_ess[F.t_offset + t] = (float)ess;
        if (a.out_loglik_t) a.out_loglik_t[F.t_offset + t] = loglik;
      }
      // swap buffers
      float* tx = xa;
      xa = xb;
      xb = tx;
      float* tl = lwa;
      lwa = lwb;
      lwb = tl;
      __syncthreads();
    }
    // final particles live in xa/lwa; expose them in slot 0 buffers for the host
    if (xa != a.x0 + F.state_offset * D) {
      for (int64_t i = tid; i < n; i += NT) {
        #pragma unroll
        for (int q = 0; q < D; ++q) (a.x0 + F.state_offset * D)[i * D + q] = xa[i * D + q];
        (a.lw0 + F.state_offset)[i] = lwa[i];
      }
    }
    if (tid == 0) {
      a.out_loglik[f] = loglik;
      if (GRAD)
        for (int k = 0; k < 3; ++k) a.out_grad[(int64_t)f * 3 + k] = gtot[k];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- standalone resampling
// Two launches: the exact prefix sums by one workgroup, then one thread per output for the searches
// (the same sums and searches as resample_block).
// from_log: w holds log-weights, the weights used are exp(w - max w) (f32).
__global__ __launch_bounds__(NT) void resample_csum_kernel(const float* __restrict__ w, int64_t n, int from_log,
                                                           uint64_t* csum) {
  __shared__ BlockScratch s;
  float m = 0.f;
  if (from_log) {
    float mm = -INFINITY;
    for (int64_t i = threadIdx.x; i < n; i += NT) mm = fmaxf(mm, w[i]);
    m = block_max(mm, s.fm);
    __syncthreads();
  }
  auto wfn = [&](int64_t i) {
    const float v = from_log ? __expf(w[i] - m) : w[i];
    return v > 0.f ? v : 0.f;
  };
  quantised_csum_block(wfn, n, csum, s);
}

__global__ __launch_bounds__(256) void resample_search_kernel(const uint64_t* __restrict__ csum, int64_t n,
                                                              const uint32_t* key, int method, int32_t* idx) {
  const Key k{key[0], key[1]};
  const double W = (double)csum[n - 1];  // the inclusive sum at the last index is the total
  const double u_sys = method == 0 ? systematic_uniform(k) : 0.0;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    idx[j] = ancestor_at(csum, n, W, k, method, u_sys, j);
}

// random.categorical(key, lw, shape=(n_out,)): one wave per output j, the lanes over the n Gumbel
// draws of row j (elements j n + i of the (n_out, n) bits array); each lane keeps its first maximum,
// and a butterfly takes the larger value, the lower index on ties (the first maximum overall).
__global__ __launch_bounds__(256) void categorical_kernel(const float* __restrict__ lw, int64_t n, int64_t n_out,
                                                          const uint32_t* key, int32_t* idx) {
  const Key k{key[0], key[1]};
  const uint32_t S = (uint32_t)(n * n_out);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t j = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; j < n_out; j += waves) {
    float best = -INFINITY;
    int64_t bi = -1;
    for (int64_t i = lane; i < n; i += kWave) {
      const float u = uniform_from_bits(bits_at(k, S, (uint32_t)(j * n + i)), 1.17549435e-38f, 1.f);
      const float g = -logf(-logf(u)) + lw[i];
      if (bi < 0 || g > best) {
        best = g;
        bi = i;
      }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const float ob = __shfl_xor(best, o, kWave);
      const int64_t oi = (int64_t)__shfl_xor((long long)bi, o, kWave);
      if (oi >= 0 && (bi < 0 || ob > best || (ob == best && oi < bi))) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) idx[j] = (int32_t)bi;
  }
}

// ---------------------------------------------------------------- large filters: grid-wide launches per step
// One workgroup per filter leaves a large filter running long after the others have finished. Filters
// flagged kPfLarge instead advance one time step per sequence of three launches over 1024-particle
// chunks (all flagged filters batched in every launch):
//   lf_prep    per chunk: combine the chunk accumulators of step t-1 in chunk order (every chunk of
//              the filter alike; its first chunk writes the log-likelihood, ESS and filter state),
//              the resampling decision and weight shift for step t, and the chunk's sum of quantised
//              weights (resampling steps only);
//   lf_csum    per chunk: the chunk's threefry blocks of the step's split(key, n) (every step), then
//              offset = sum of the earlier chunks' totals, plus an inclusive block scan;
//   lf_step    per chunk: ancestor (binary search in the exact prefix sums, or Gumbel-max), then
//              propagate, weight, outputs and the chunk accumulators of step t+1.
// The weights are shifted by float(lse) and quantised with the shift of the largest weight exactly as
// in pf_kernel, so for the same log-weights and key the ancestors are the same.
constexpr int LT = 256;          // threads per workgroup
constexpr int LW = LT / kWave;
constexpr int PPT = 4;           // particles per thread
constexpr int CH = LT * PPT;     // particles per chunk (one workgroup)
constexpr int SG = 256;          // prefix-sum search granularity (chunk ends staged in LDS)
constexpr int kSearchWin = 2048;   // prefix sums of a block's ancestor chunks staged in LDS (16 KB)
constexpr int kChunkEnds = 4096;   // SG-block end sums staged in LDS (filters up to 2^20 particles);
                                   // dynamic LDS sized by the launch's largest filter (at most 32 KB)

struct LfDesc {
  int64_t n, T, y_offset, state_offset, out_offset, t_offset, chunk0;
  int64_t key_slot;  // this filter's T step-key slots in LfArgs::skeys (disjoint per filter)
  uint32_t key[2];
  int32_t param_offset, f, flags;
  int32_t slot;  // index of the filter in desc and st
};

struct LfState {
  double lse_res;  // lse of the weights the next increments add to (log n after resampling)
  double loglik;
  double gprev[3], gtot[3];
  float m;    // weight shift, float(lse of the last step)
  int res;    // resample at the next step
  int shift;  // quantisation shift of the next resampling
  int pad;
};

// Per filter and step t >= 1: the step key (split(key, T-1)[t-1], or the filter key with
// kPfStepKey), the optimal proposal's first split key split(step key, n)[0], and the systematic
// uniform; derived for all steps at once by the init launch, read by the later launches.
struct StepKey {
  uint32_t k[4];
  float su;
  uint32_t pad;
};

struct LfArgs {
  const LfDesc* desc;
  const LfDesc* cdesc;  // [n_chunks] the descriptor of each chunk's filter (one load per workgroup)
  int n_large;
  int64_t n_chunks;
  const float* params;
  const float* y;
  const float* step_scale;
  float* x0;
  float* x1;
  float* lw0;
  float* lw1;
  uint64_t* csum;
  float* tan0;
  float* tan1;
  int resampling;
  float ess_threshold;
  float* out_value;
  float* out_lw;
  int32_t* out_anc;
  float* out_ess;
  double* out_loglik_t;
  double* out_loglik;
  double* out_grad;
  float* part_m;     // [n_chunks] chunk LSE accumulators (max, sums in FP64)
  double* part_s1;   // [n_chunks]
  double* part_s2;   // [n_chunks]
  double* part_g;    // [n_chunks][3] tangent sums (GRAD)
  uint64_t* ctot;    // [n_chunks] quantised weight totals
  LfState* st;       // [n_large]
  uint64_t* cends;   // [n_chunks * CH / SG] prefix sums at the SG-block ends, per filter from chunk0 * CH / SG
  int64_t* bounds;   // [n_chunks] systematic: ancestor of each chunk's first target
  StepKey* skeys;    // [key_slot + t] per filter
  uint32_t* bar;     // persistent kernel: [0] barrier arrivals, [1] abort flag (zeroed by the init launch)
  uint32_t* pkeys;   // [n_chunks * CH * 2] the step's split(step key, n), flat, per filter from chunk0 * CH * 2;
                     // null for the optimal proposal (its noise is not drawn from per-particle keys)
};

struct LfScratch {
  float fm[LW];
  double d1[LW], d2[LW], g[LW][3];
  uint64_t u[LW];
};

// chunk-level reductions in a fixed order (wave butterflies, then waves in order)
MOCAT_DEV uint64_t lf_block_sum(uint64_t v, LfScratch& s) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) s.u[w] = v;
  __syncthreads();
  uint64_t r = 0;
  for (int i = 0; i < LW; ++i) r += s.u[i];
  return r;
}

MOCAT_DEV uint64_t lf_block_inclusive_scan(uint64_t v, LfScratch& s) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t incl = wave_inclusive_scan(v);
  __syncthreads();
  if (lane == kWave - 1) s.u[w] = incl;
  __syncthreads();
  uint64_t off = 0;
  for (int i = 0; i < w; ++i) off += s.u[i];
  return off + incl;
}

// Chunk accumulator in two passes, one exponential per value: the workgroup maximum of the values
// (NaN excluded; a NaN anywhere makes the result's m NaN), then sum exp(v - M) and its square in FP64,
// lanes by butterflies and waves in order. Every thread gets the result.
template <int P>
MOCAT_DEV LseAcc lf_block_lse2(const float (&v)[P], const bool (&live)[P], LfScratch& s) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  float m = -INFINITY;
  int bad = 0;
#pragma unroll
  for (int pk = 0; pk < P; ++pk) {
    if (!live[pk]) continue;
    if (isnan(v[pk]))
      bad = 1;
    else
      m = fmaxf(m, v[pk]);
  }
  m = wave_max(m);
  bad = __any(bad) ? 1 : 0;
  __syncthreads();
  if (lane == 0) {
    s.fm[w] = m;
    s.u[w] = (uint64_t)bad;
  }
  __syncthreads();
  float M = s.fm[0];
  bool anybad = s.u[0] != 0;
  for (int i = 1; i < LW; ++i) {
    M = fmaxf(M, s.fm[i]);
    anybad = anybad || s.u[i] != 0;
  }
  double s1 = 0.0, s2 = 0.0;
  if (M != -INFINITY) {
#pragma unroll
    for (int pk = 0; pk < P; ++pk) {
      if (!live[pk] || !(v[pk] > -INFINITY)) continue;  // -inf and NaN add nothing
      const double e = exp((double)v[pk] - (double)M);
      s1 += e;
      s2 += e * e;
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  __syncthreads();
  if (lane == 0) {
    s.d1[w] = s1;
    s.d2[w] = s2;
  }
  __syncthreads();
  double r1 = s.d1[0], r2 = s.d2[0];
  for (int i = 1; i < LW; ++i) {
    r1 += s.d1[i];
    r2 += s.d2[i];
  }
  return LseAcc{anybad ? NAN : M, r1, r2};
}

// As lf_block_lse2, also summing exp(v - M) * t[q] for three tangents (returned in g, relative to M).
template <int P>
MOCAT_DEV LseAcc lf_block_lse2g(const float (&v)[P], const float (&t)[P][3], const bool (&live)[P], LfScratch& s,
                                double (&g)[3]) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  float m = -INFINITY;
  int bad = 0;
#pragma unroll
  for (int pk = 0; pk < P; ++pk) {
    if (!live[pk]) continue;
    if (isnan(v[pk]))
      bad = 1;
    else
      m = fmaxf(m, v[pk]);
  }
  m = wave_max(m);
  bad = __any(bad) ? 1 : 0;
  __syncthreads();
  if (lane == 0) {
    s.fm[w] = m;
    s.u[w] = (uint64_t)bad;
  }
  __syncthreads();
  float M = s.fm[0];
  bool anybad = s.u[0] != 0;
  for (int i = 1; i < LW; ++i) {
    M = fmaxf(M, s.fm[i]);
    anybad = anybad || s.u[i] != 0;
  }
  double s1 = 0.0, s2 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (M != -INFINITY) {
#pragma unroll
    for (int pk = 0; pk < P; ++pk) {
      if (!live[pk] || !(v[pk] > -INFINITY)) continue;
      const double e = exp((double)v[pk] - (double)M);
      s1 += e;
      s2 += e * e;
      g0 += e * t[pk][0];
      g1 += e * t[pk][1];
      g2 += e * t[pk][2];
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  g0 = wave_sum(g0);
  g1 = wave_sum(g1);
  g2 = wave_sum(g2);
  __syncthreads();
  if (lane == 0) {
    s.d1[w] = s1;
    s.d2[w] = s2;
    s.g[w][0] = g0;
    s.g[w][1] = g1;
    s.g[w][2] = g2;
  }
  __syncthreads();
  double r1 = s.d1[0], r2 = s.d2[0];
  g[0] = s.g[0][0];
  g[1] = s.g[0][1];
  g[2] = s.g[0][2];
  for (int i = 1; i < LW; ++i) {
    r1 += s.d1[i];
    r2 += s.d2[i];
    for (int q = 0; q < 3; ++q) g[q] += s.g[i][q];
  }
  return LseAcc{anybad ? NAN : M, r1, r2};
}

MOCAT_DEV void lf_store_lse_g(const LfArgs& a, int64_t c, LseAcc r, const double (&g)[3]) {
  if (threadIdx.x == 0) {
    a.part_m[c] = r.m;
    a.part_s1[c] = r.s1;
    a.part_s2[c] = r.s2;
    for (int q = 0; q < 3; ++q) a.part_g[c * 3 + q] = g[q];
  }
}

MOCAT_DEV void lf_store_lse(const LfArgs& a, int64_t c, LseAcc r) {
  if (threadIdx.x == 0) {
    a.part_m[c] = r.m;
    a.part_s1[c] = r.s1;
    a.part_s2[c] = r.s2;
  }
}

template <class Model, int D, int DY, bool GRAD>
MOCAT_DEV void lf_init_chunk(const LfArgs& a, int64_t c, LfScratch& s, Model& model) {
  const LfDesc F = a.cdesc[c];
  const int tid = threadIdx.x;
  if (a.bar && c == 0 && tid < 2) a.bar[tid] = 0u;  // the persistent kernel's barrier counter and abort flag
  load_shared_params(model, a.params + F.param_offset);
  __syncthreads();
  const int64_t n = F.n;
  const Key key{F.key[0], F.key[1]};
  const Key sub = (F.flags & kPfInitKey) ? key : split_at(key, 2, 1);
  constexpr uint32_t hD = (D + 1) / 2;
  const Key sub0 = Model::kOptimal ? split_at(sub, (uint32_t)n, 0u) : Key{0u, 0u};
  if (c == F.chunk0) {
    const Key rk = (F.flags & kPfResume) ? key : split_at(key, 2, 0);
    for (int64_t t = 1 + tid; t < F.T; t += LT) {
      const Key ikk = (F.flags & kPfStepKey) ? key : split_at(rk, (uint32_t)(F.T - 1), (uint32_t)(t - 1));
      StepKey sk;
      sk.k[0] = ikk.a;
      sk.k[1] = ikk.b;
      sk.k[2] = sk.k[3] = 0u;
      if constexpr (Model::kOptimal) {
        const Key k0 = split_at(ikk, (uint32_t)n, 0u);
        sk.k[2] = k0.a;
        sk.k[3] = k0.b;
      }
      sk.su = uniform_from_bits(bits_at(ikk, 1, 0), 0.f, 1.f);
      sk.pad = 0u;
      a.skeys[F.key_slot + t] = sk;
    }
  }
  float lv[PPT], lt[PPT][3];
  bool lvalid[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    lv[pk] = 0.f;
    lt[pk][0] = lt[pk][1] = lt[pk][2] = 0.f;
    lvalid[pk] = false;
  }
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = (c - F.chunk0) * CH + pk * LT + tid;
    if (i >= n) break;
    float eps[D], x[D];
    float l;
    if (F.flags & kPfResume) {
#pragma unroll
      for (int q = 0; q < D; ++q) x[q] = a.x0[(F.state_offset + i) * D + q];
      l = a.lw0[F.state_offset + i];
    } else if constexpr (Model::kOptimal) {
      shared_eps<D>(sub0, (uint32_t)(n * D), (uint32_t)i, eps);
      model.init_opt(eps, a.y + F.y_offset, x);
      l = 0.f;
    } else if constexpr (GRAD) {
      const Key ki = split_at(sub, (uint32_t)n, (uint32_t)i);
#pragma unroll
      for (uint32_t q = 0; q < hD; ++q) {
        uint32_t w0, w1;
        bits_pair(ki, D, q, w0, w1);
        eps[q] = normal_from_bits(w0);
        if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
      }
      model.init(eps, x);
      float dldx, dldb;
      l = -model.lik_grad(x, a.y + F.y_offset, dldx, dldb);
      float* ta = a.tan0 + F.state_offset * 5;
      ta[i * 5 + 0] = 0.f;
      ta[i * 5 + 1] = 0.f;
      ta[i * 5 + 2] = 0.f;
      ta[i * 5 + 3] = 0.f;
      ta[i * 5 + 4] = -dldb;
      lt[pk][2] = -dldb;
    } else {
      const Key ki = split_at(sub, (uint32_t)n, (uint32_t)i);
#pragma unroll
      for (uint32_t q = 0; q < hD; ++q) {
        uint32_t w0, w1;
        bits_pair(ki, D, q, w0, w1);
        eps[q] = normal_from_bits(w0);
        if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
      }
      model.init(eps, x);
      l = -model.lik(x, a.y + F.y_offset);
    }
    float* xa = a.x0 + F.state_offset * D;
#pragma unroll
    for (int q = 0; q < D; ++q) xa[i * D + q] = x[q];
    a.lw0[F.state_offset + i] = l;
    lv[pk] = l;
    lvalid[pk] = true;
    if (F.out_offset >= 0) {
      if (a.out_value)
#pragma unroll
        for (int q = 0; q < D; ++q) a.out_value[(F.out_offset + i) * D + q] = x[q];
      if (a.out_lw) a.out_lw[F.out_offset + i] = l;
      if (a.out_anc) a.out_anc[F.out_offset + i] = (int32_t)i;
    }
  }
  if constexpr (GRAD) {
    double g[3];
    const LseAcc r = lf_block_lse2g<PPT>(lv, lt, lvalid, s, g);
    lf_store_lse_g(a, c, r, g);
  } else {
    lf_store_lse(a, c, lf_block_lse2<PPT>(lv, lvalid, s));
  }
}

template <class Model, int D, int DY, bool GRAD>
__global__ __launch_bounds__(LT) void lf_init_kernel(LfArgs a) {
  __shared__ LfScratch s;
  __shared__ Model model;
  lf_init_chunk<Model, D, DY, GRAD>(a, blockIdx.x, s, model);
}

// The prefix sums are integers below 2^52 (exact in FP64), so for a target T >= 0 the test
// (double)C <= T is C <= floor(T): targets are converted once and every level compares integers.
MOCAT_DEV uint64_t target_floor(double t) { return t > 0.0 ? (uint64_t)floor(t) : 0ull; }

MOCAT_DEV const float* lf_lw_prev(const LfArgs& a, int64_t t) { return ((t - 1) & 1) ? a.lw1 : a.lw0; }

// One launch per step t >= 1 (and one after the last step):
//  * the statistics of step t - 1 (LSE of its weights, ESS, log-likelihood, next resampling decision
//    and weight shifts), reduced from the chunk partials in a fixed order by every chunk of the filter
//    (identical results in every workgroup); the filter's first chunk also updates the filter state
//    and writes the per-step outputs. The decision and shifts do not depend on the previous state, so
//    no workgroup reads what another writes in this launch;
//  * for step t, when it resamples, the chunk's total of quantised weights.
template <bool GRAD>
MOCAT_DEV void lf_prep_chunk(const LfArgs& a, int64_t t, int64_t c, LfScratch& s) {
  const LfDesc F = a.cdesc[c];
  const int fi = F.slot;
  if (t - 1 >= F.T) return;
  const int64_t n = F.n, nchunk = (n + CH - 1) / CH;
  // the chunk's weights for its quantised total are loaded first: they do not depend on the combine
  const bool want_tot = t < F.T && a.resampling != 2;
  const float* lwp = lf_lw_prev(a, t) + F.state_offset;
  float lwv[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = (c - F.chunk0) * CH + pk * LT + threadIdx.x;
    lwv[pk] = (want_tot && i < n) ? lwp[i] : 0.f;
  }
  double gsum[3] = {0.0, 0.0, 0.0};  // tangent sums relative to the filter's maximum (GRAD)
  LseAcc r;
  {
    // two passes over the chunk partials: their maximum, then the rescaled sums (one exponential each;
    // the tangent sums of a chunk are relative to its maximum, like its weight sums). Filters of at
    // most LT chunks (one partial per thread) load each partial once, in the first pass.
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const bool one = nchunk <= LT;
    float m = -INFINITY, pm1 = -INFINITY;
    double s1v = 0.0, s2v = 0.0, gv[3] = {0.0, 0.0, 0.0};
    int bad = 0;
    if (one) {
      if (threadIdx.x < nchunk) {
        const int64_t cc = F.chunk0 + threadIdx.x;
        pm1 = a.part_m[cc];
        s1v = a.part_s1[cc];
        s2v = a.part_s2[cc];
        if constexpr (GRAD)
          for (int q = 0; q < 3; ++q) gv[q] = a.part_g[cc * 3 + q];
        if (isnan(pm1))
          bad = 1;
        else
          m = pm1;
      }
    } else {
      for (int64_t k = threadIdx.x; k < nchunk; k += LT) {
        const float pm = a.part_m[F.chunk0 + k];
        if (isnan(pm))
          bad = 1;
        else
          m = fmaxf(m, pm);
      }
    }
    m = wave_max(m);
    bad = __any(bad) ? 1 : 0;
    __syncthreads();
    if (lane == 0) {
      s.fm[w] = m;
      s.u[w] = (uint64_t)bad;
    }
    __syncthreads();
    float M = s.fm[0];
    bool anybad = s.u[0] != 0;
    for (int i = 1; i < LW; ++i) {
      M = fmaxf(M, s.fm[i]);
      anybad = anybad || s.u[i] != 0;
    }
    double s1 = 0.0, s2 = 0.0, g[3] = {0.0, 0.0, 0.0};
    if (M != -INFINITY && one) {
      if (pm1 > -INFINITY) {
        const double f = exp((double)pm1 - (double)M);
        s1 += s1v * f;
        s2 += s2v * f * f;
        if constexpr (GRAD)
          for (int q = 0; q < 3; ++q) g[q] += gv[q] * f;
      }
    } else if (M != -INFINITY) {
      for (int64_t k = threadIdx.x; k < nchunk; k += LT) {
        const int64_t cc = F.chunk0 + k;
        const float pm = a.part_m[cc];
        if (!(pm > -INFINITY)) continue;
        const double f = exp((double)pm - (double)M);
        s1 += a.part_s1[cc] * f;
        s2 += a.part_s2[cc] * f * f;
        if constexpr (GRAD)
          for (int q = 0; q < 3; ++q) g[q] += a.part_g[cc * 3 + q] * f;
      }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if constexpr (GRAD)
      for (int q = 0; q < 3; ++q) g[q] = wave_sum(g[q]);
    __syncthreads();
    if (lane == 0) {
      s.d1[w] = s1;
      s.d2[w] = s2;
      if constexpr (GRAD)
        for (int q = 0; q < 3; ++q) s.g[w][q] = g[q];
    }
    __syncthreads();
    double r1 = s.d1[0], r2 = s.d2[0];
    if constexpr (GRAD)
      for (int q = 0; q < 3; ++q) gsum[q] = s.g[0][q];
    for (int i = 1; i < LW; ++i) {
      r1 += s.d1[i];
      r2 += s.d2[i];
      if constexpr (GRAD)
        for (int q = 0; q < 3; ++q) gsum[q] += s.g[i][q];
    }
    r = LseAcc{anybad ? NAN : M, r1, r2};
  }
  double lse, lse2;
  if (isnan(r.m)) {
    lse = NAN;
    lse2 = NAN;
  } else if (r.m == -INFINITY) {
    lse = -INFINITY;
    lse2 = -INFINITY;
  } else {
    lse = (double)r.m + log(r.s1);
    lse2 = 2.0 * (double)r.m + log(r.s2);
  }
  const double logn = log((double)n);
  const double ess = exp(2.0 * lse - lse2);
  // decision and weight shift for step t (pf_kernel: m = float(lse), shift from the largest weight
  // exp(lw - m), i.e. exp(max lw - m): the hardware exponential is monotone)
  const int res = ess < (double)a.ess_threshold * (double)n ? 1 : 0;
  const float m = (float)lse;
  const int shift = quant_shift(__expf(r.m - m), n);
  if (c == F.chunk0) {
    double gm[3] = {0.0, 0.0, 0.0};
    if constexpr (GRAD)
      for (int q = 0; q < 3; ++q) gm[q] = gsum[q] / r.s1;
    if (threadIdx.x == 0) {
      const int64_t tp = t - 1;
      LfState st = a.st[fi];
      st.loglik = tp == 0 ? ((F.flags & kPfResume) ? 0.0 : lse - logn) : st.loglik + lse - st.lse_res;
      if (a.out_ess) a.out_ess[F.t_offset + tp] = (float)ess;
      if (a.out_loglik_t) a.out_loglik_t[F.t_offset + tp] = st.loglik;
      if constexpr (GRAD) {
        for (int q = 0; q < 3; ++q) {
          st.gtot[q] = tp == 0 ? gm[q] : st.gtot[q] + gm[q] - (st.res ? 0.0 : st.gprev[q]);
          st.gprev[q] = gm[q];
        }
      }
      st.res = res;
      st.lse_res = res ? logn : lse;
      st.m = m;
      st.shift = shift;
      a.st[fi] = st;
      if (tp == F.T - 1) {
        a.out_loglik[F.f] = st.loglik;
        if constexpr (GRAD)
          for (int q = 0; q < 3; ++q) a.out_grad[(int64_t)F.f * 3 + q] = st.gtot[q];
      }
    }
  }
  if (t >= F.T || !res || a.resampling == 2) return;
  uint64_t q = 0;
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = (c - F.chunk0) * CH + pk * LT + threadIdx.x;
    if (i < n) q += quantise(__expf(lwv[pk] - m), shift);
  }
  const uint64_t tot = lf_block_sum(q, s);
  if (threadIdx.x == 0) a.ctot[c] = tot;
}

template <bool GRAD>
__global__ __launch_bounds__(LT) void lf_prep_kernel(LfArgs a, int64_t t) {
  __shared__ LfScratch s;
  lf_prep_chunk<GRAD>(a, t, blockIdx.x, s);
}

// cs: CH entries of LDS (this chunk's prefix sums, for the systematic boundary searches)
MOCAT_DEV void lf_csum_chunk(const LfArgs& a, int64_t t, int64_t c, LfScratch& s, uint64_t* cs) {
  const LfDesc F = a.cdesc[c];
  const int fi = F.slot;
  if (t >= F.T) return;
  const int64_t local = c - F.chunk0;
  const int64_t nchunk = (F.n + CH - 1) / CH;
  // loads that do not depend on each other are issued first: the filter state, the chunk's weights and
  // (one per thread, for filters of at most LT chunks) the chunk totals; the weights and totals are
  // used on resampling steps only
  const LfState st = a.st[fi];
  const bool sums = a.resampling != 2;
  const int64_t i0 = local * CH + (int64_t)threadIdx.x * PPT;  // thread tid owns PPT consecutive particles
  const float* lw = lf_lw_prev(a, t) + F.state_offset;
  float lwv[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) lwv[pk] = (sums && i0 + pk < F.n) ? lw[i0 + pk] : 0.f;
  const bool one = nchunk <= LT;
  const uint64_t tot1 = (sums && one && threadIdx.x < nchunk) ? a.ctot[F.chunk0 + threadIdx.x] : 0ull;
  if (a.pkeys) {
    // split(step key, n) for the step kernel, one threefry block per key: block j of split's counter
    // pairs (j, n + j) yields words j and n + j of the flat key array (key i = words 2i, 2i + 1), so
    // the chunk's blocks j in [local * CH, (local + 1) * CH) fill words in both halves
    const StepKey skv = a.skeys[F.key_slot + t];
    const uint32_t un = (uint32_t)F.n;
    uint32_t* PK = a.pkeys + F.chunk0 * (2 * CH);
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk) {
      const int64_t j = local * CH + pk * LT + threadIdx.x;
      if (j < F.n) {
        uint32_t x0 = (uint32_t)j, x1 = (uint32_t)j + un;
        threefry2x32(skv.k[0], skv.k[1], x0, x1);
        PK[j] = x0;
        PK[j + F.n] = x1;
      }
    }
  }
  if (!st.res || !sums) return;
  uint64_t off = 0, wsum = 0;
  if (one) {
    wsum = tot1;
    off = threadIdx.x < local ? tot1 : 0ull;
  } else {
    for (int64_t k = threadIdx.x; k < nchunk; k += LT) {
      const uint64_t v = a.ctot[F.chunk0 + k];
      wsum += v;
      if (k < local) off += v;
    }
  }
  off = lf_block_sum(off, s);
  wsum = lf_block_sum(wsum, s);
  const float su = a.skeys[F.key_slot + t].su;  // the step's systematic uniform (init launch)
  uint64_t q[PPT], own = 0;
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    q[pk] = i0 + pk < F.n ? quantise(__expf(lwv[pk] - st.m), st.shift) : 0ull;
    own += q[pk];
  }
  uint64_t run = off + lf_block_inclusive_scan(own, s) - own;
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    run += q[pk];
    const int64_t i = i0 + pk;
    cs[threadIdx.x * PPT + pk] = run;
    if (i < F.n) {
      a.csum[F.state_offset + i] = run;
      // the SG-block end sums, contiguous per filter, for the step kernel's first search level
      if ((i + 1) % SG == 0 || i == F.n - 1) a.cends[F.chunk0 * (CH / SG) + i / SG] = run;
    }
  }
  if (a.resampling != 0) return;
  // Systematic targets floor((j + u) W / n) rise with j, so chunk b's ancestors lie between the
  // ancestors of its first target and of the next chunk's first target. The chunk whose sums hold a
  // boundary target (earlier sums <= target < its last sum; the last chunk also takes targets at or
  // beyond the total, clamped to n - 1 as in the search) finds its ancestor in LDS and records it.
  __syncthreads();
  const int64_t count = (F.n - local * CH) < CH ? (F.n - local * CH) : CH;
  const uint64_t chunk_hi = cs[count - 1];
  const bool last = local == nchunk - 1;
  const double sstep = (double)wsum / (double)F.n;
  for (int64_t b = threadIdx.x; b < nchunk; b += LT) {
    const uint64_t tg = target_floor(((double)(b * CH) + (double)su) * sstep);
    if (tg < off || (!last && tg >= chunk_hi)) continue;
    int64_t lo = 0, len = count;
    while (len > 0) {
      const int64_t h = len >> 1;
      if (cs[lo + h] <= tg) {
        lo += h + 1;
        len -= h + 1;
      } else {
        len = h;
      }
    }
    a.bounds[F.chunk0 + b] = local * CH + (lo < count ? lo : count - 1);
  }
}

__global__ __launch_bounds__(LT) void lf_csum_kernel(LfArgs a, int64_t t) {
  __shared__ LfScratch s;
  __shared__ uint64_t cs[CH];
  lf_csum_chunk(a, t, blockIdx.x, s, cs);
}

// dst[k] = src(k) for k < count, by the whole workgroup, U loads per thread issued before their stores.
template <int U, class Src>
MOCAT_DEV void stage_lds(uint64_t* dst, int64_t count, Src src) {
  for (int64_t k0 = 0; k0 < count; k0 += (int64_t)U * LT) {
    uint64_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + (int64_t)u * LT + threadIdx.x;
      v[u] = k < count ? src(k) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + (int64_t)u * LT + threadIdx.x;
      if (k < count) dst[k] = v[u];
    }
  }
}

// For P searches at once: out[p] = the first k in [b[p], b[p] + m[p]) with v(k) > tg[p] (v
// non-decreasing), or b[p] + m[p] when there is none; inactive searches return b[p]. Each halving
// step issues the reads of all unfinished searches before comparing.
template <int P, class Get>
MOCAT_DEV void upper_bound_lockstep(Get v, const int64_t* b, const int64_t* m, const uint64_t* tg, const bool* act,
                                    int64_t* out) {
  int64_t lo[P], len[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    lo[q] = b[q];
    len[q] = act[q] ? m[q] : 0;
  }
  while (true) {
    bool any = false;
#pragma unroll
    for (int q = 0; q < P; ++q) any |= len[q] > 0;
    if (!any) break;
    uint64_t val[P];
#pragma unroll
    for (int q = 0; q < P; ++q) val[q] = len[q] > 0 ? v(lo[q] + (len[q] >> 1)) : 0ull;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      if (len[q] <= 0) continue;
      const int64_t h = len[q] >> 1;
      if (val[q] <= tg[q]) {
        lo[q] += h + 1;
        len[q] -= h + 1;
      } else {
        len[q] = h;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < P; ++q) out[q] = lo[q];
}

// win: kSearchWin entries of LDS; cend: cend_slots entries (staged by the multinomial search only)
// For P targets that do not decrease with p (a thread's consecutive systematic targets): out[p] = the
// first k in [b, e) with v(k) > tg[p] (v non-decreasing), or e; inactive targets (the chunk's tail, all
// after the active ones) return the previous result. The first target takes a binary search; each
// next one steps forward from the previous result (its answer is no smaller) up to four times
// (offspring counts average one) and finishes with a binary search over the rest when those do not
// reach it.
template <int P, class Get>
MOCAT_DEV void upper_bound_rising(Get v, int64_t b, int64_t e, const uint64_t* tg, const bool* act, int64_t* out) {
  int64_t lo = b;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (!act[q]) {
      out[q] = lo;
      continue;
    }
    if (q > 0)
      for (int step = 0; step < 4 && lo < e && v(lo) <= tg[q]; ++step) ++lo;
    int64_t len = e - lo;
    if (q > 0 && (len <= 0 || v(lo) > tg[q])) len = 0;  // found by the steps
    while (len > 0) {
      const int64_t h = len >> 1;
      if (v(lo + h) <= tg[q]) {
        lo += h + 1;
        len -= h + 1;
      } else {
        len = h;
      }
    }
    out[q] = lo;
  }
}

template <class Model, int D, int DY, bool GRAD>
MOCAT_DEV void lf_step_chunk(const LfArgs& a, int64_t t, int cend_slots, int64_t c, LfScratch& s, Model& model,
                             uint64_t* win, uint64_t* cend) {
  const LfDesc F = a.cdesc[c];
  const int fi = F.slot;
  if (t >= F.T) return;
  const int tid = threadIdx.x;
  const LfState st = a.st[fi];
  const int64_t n = F.n, base = (c - F.chunk0) * CH;
  const bool res = st.res != 0;
  // particle of slot pk: four consecutive particles per thread (systematic targets rise within the
  // thread: one search and forward steps), or, with the gradient, particles LT apart (the tangent
  // records of 5 floats per particle are then stored by consecutive lanes)
  constexpr bool kConsecutive = !GRAD;
  auto pidx = [&](int pk) -> int64_t { return kConsecutive ? base + tid * PPT + pk : base + pk * LT + tid; };
  const bool odd = (t - 1) & 1;
  const float* xa = (odd ? a.x1 : a.x0) + F.state_offset * D;
  float* xb = (odd ? a.x0 : a.x1) + F.state_offset * D;
  const float* lwa = (odd ? a.lw1 : a.lw0) + F.state_offset;
  float* lwb = (odd ? a.lw0 : a.lw1) + F.state_offset;
  const float* ta = GRAD ? (odd ? a.tan1 : a.tan0) + F.state_offset * 5 : nullptr;
  float* tb = GRAD ? (odd ? a.tan0 : a.tan1) + F.state_offset * 5 : nullptr;
  const float* ycur = a.y + F.y_offset + t * DY;
  constexpr uint32_t hD = (D + 1) / 2;

  // Ancestors by a two-level search in the exact prefix sums, without long chains of dependent
  // global loads: the last sum of every SG-particle block is staged in LDS and searched first, then
  // the first index inside the block found. Systematic targets rise with the particle index, so this
  // chunk's ancestors lie in a contiguous run of blocks, staged in LDS too when it is short. Each
  // thread searches for its PPT particles in lockstep (independent LDS or global reads overlap).
  // (res and the method are per filter: the branches are uniform over the workgroup.)
  const uint64_t* C = a.csum + F.state_offset;
  const int64_t nsg = (n + SG - 1) / SG;
  // systematic: the chunk's ancestor window comes from the prefix-sum pass (bounds); multinomial
  // targets are unordered and take the two-level search over the staged block-end sums
  const bool window = res && a.resampling == 0;
  const bool two_level = res && a.resampling == 1 && nsg <= cend_slots;
  bool staged = false;
  int64_t lo = 0, whi = n - 1;
  if (window) {
    const int64_t cl = c - F.chunk0, nchunk = (n + CH - 1) / CH;
    const int64_t b0 = a.bounds[c], b1 = cl + 1 < nchunk ? a.bounds[c + 1] : n - 1;
    lo = b0 < 0 ? 0 : (b0 > n - 1 ? n - 1 : b0);
    whi = b1 < lo ? lo : (b1 > n - 1 ? n - 1 : b1);
  }
  const double W = (res && a.resampling != 2) ? (double)C[n - 1] : 0.0;
  const double sstep = W / (double)n;
  // the chunk-end sums go first: their loads are in flight while thread 0 derives the step key
  const uint64_t* CE = a.cends + F.chunk0 * (CH / SG);
  if (two_level) stage_lds<8>(cend, nsg, [&](int64_t k) { return CE[k]; });
  // the step key, the optimal proposal's first split key and the systematic uniform (init launch)
  const StepKey skv = a.skeys[F.key_slot + t];
  const uint32_t sk[4] = {skv.k[0], skv.k[1], skv.k[2], skv.k[3]};
  const float su = skv.su;
  load_shared_params(model, a.params + F.param_offset);
  // the particles' split keys (lf_csum) are loaded before the ancestor search, which they do not need
  const uint32_t* PK = a.pkeys + F.chunk0 * (2 * CH);
  uint2 kwp[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = pidx(pk);
    kwp[pk] = (!Model::kOptimal && i < n) ? *reinterpret_cast<const uint2*>(PK + 2 * i) : make_uint2(0u, 0u);
  }
  __syncthreads();
  const Key ik{sk[0], sk[1]};
  auto cend_at = [&](int64_t k) { return cend[k]; };
  if (window && whi - lo + 1 <= kSearchWin) {
    stage_lds<8>(win, whi - lo + 1, [&](int64_t k) { return C[lo + k]; });
    __syncthreads();
    staged = true;
  }

  // ancestors of the thread's PPT particles
  int64_t srcs[PPT];
  bool live[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    srcs[pk] = pidx(pk);
    live[pk] = srcs[pk] < n;
  }
  if (res && a.resampling == 2) {
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk)
      if (live[pk]) srcs[pk] = categorical_row(lwa, n, ik, (uint32_t)(n * n), srcs[pk]);
  } else if (res) {
    uint64_t tg[PPT];
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk) {
      const int64_t i = srcs[pk];
      tg[pk] = target_floor(a.resampling == 0
                                ? ((double)i + (double)su) * sstep
                                : (double)uniform_from_bits(bits_at(ik, (uint32_t)n, (uint32_t)i), 0.f, 1.f) * W);
    }
    if (window) {
      // the first index of [lo, whi] whose sum exceeds the target (whi + 1: none does, clamped)
      int64_t p[PPT];
      if constexpr (kConsecutive) {
        if (staged)
          upper_bound_rising<PPT>([&](int64_t k) { return win[k - lo]; }, lo, whi + 1, tg, live, p);
        else
          upper_bound_rising<PPT>([&](int64_t k) { return C[k]; }, lo, whi + 1, tg, live, p);
      } else {
        int64_t b[PPT], m[PPT];
#pragma unroll
        for (int pk = 0; pk < PPT; ++pk) {
          b[pk] = lo;
          m[pk] = whi - lo + 1;
        }
        if (staged)
          upper_bound_lockstep<PPT>([&](int64_t k) { return win[k - lo]; }, b, m, tg, live, p);
        else
          upper_bound_lockstep<PPT>([&](int64_t k) { return C[k]; }, b, m, tg, live, p);
      }
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk)
        if (live[pk]) srcs[pk] = p[pk] < n ? p[pk] : n - 1;
    } else if (two_level) {
      // first SG block whose last sum exceeds the target (nsg: none does), then the first index in it
      int64_t kc[PPT], b[PPT], m[PPT], zero[PPT], all[PPT];
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk) {
        zero[pk] = 0;
        all[pk] = nsg;
      }
      upper_bound_lockstep<PPT>(cend_at, zero, all, tg, live, kc);
      bool in[PPT];
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk) {
        in[pk] = live[pk] && kc[pk] < nsg;
        b[pk] = kc[pk] * SG;
        // the block's last index has a sum above the target: search the indices before it
        m[pk] = in[pk] ? ((kc[pk] + 1) * SG < n ? (kc[pk] + 1) * SG : n) - 1 - b[pk] : 0;
      }
      int64_t p[PPT];
      if (staged)
        upper_bound_lockstep<PPT>([&](int64_t k) { return win[k - lo]; }, b, m, tg, in, p);
      else
        upper_bound_lockstep<PPT>([&](int64_t k) { return C[k]; }, b, m, tg, in, p);
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk)
        if (live[pk]) srcs[pk] = in[pk] ? p[pk] : n - 1;
    } else {
      int64_t p[PPT], zero[PPT], all[PPT];
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk) {
        zero[pk] = 0;
        all[pk] = n;
      }
      upper_bound_lockstep<PPT>([&](int64_t k) { return C[k]; }, zero, all, tg, live, p);
#pragma unroll
      for (int pk = 0; pk < PPT; ++pk)
        if (live[pk]) srcs[pk] = p[pk] < n ? p[pk] : n - 1;
    }
  }
  // ancestor states of all PPT particles loaded together
  float xpre[PPT][D];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk)
#pragma unroll
    for (int q = 0; q < D; ++q) xpre[pk][q] = live[pk] ? xa[srcs[pk] * D + q] : 0.f;

  // the PPT particles in lockstep, phase by phase, so that their independent threefry chains, loads
  // and arithmetic interleave; the accumulators still take them in particle order
  float eps[PPT][D], xn[PPT][D], lwp[PPT];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = pidx(pk);
    lwp[pk] = (live[pk] && !res) ? lwa[i] : 0.f;
    if (!live[pk]) continue;
    if constexpr (Model::kOptimal) {
      shared_eps<D>(Key{sk[2], sk[3]}, (uint32_t)(n * D), (uint32_t)i, eps[pk]);
      continue;
    }
    const Key ki{kwp[pk].x, kwp[pk].y};  // split(step key, n)[i] (lf_csum)
#pragma unroll
    for (uint32_t q = 0; q < hD; ++q) {
      uint32_t w0, w1;
      bits_pair(ki, D, q, w0, w1);
      eps[pk][q] = normal_from_bits(w0);
      if (q + hD < (uint32_t)D) eps[pk][q + hD] = normal_from_bits(w1);
    }
  }
  float tg[PPT][5];
  if constexpr (GRAD) {
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk) {
      const int64_t i = pidx(pk);
      tg[pk][0] = live[pk] ? ta[srcs[pk] * 5 + 0] : 0.f;
      tg[pk][1] = live[pk] ? ta[srcs[pk] * 5 + 1] : 0.f;
#pragma unroll
      for (int q = 2; q < 5; ++q) tg[pk][q] = (live[pk] && !res) ? ta[i * 5 + q] : 0.f;
    }
  }
  float l[PPT], gd[PPT][3];
  if constexpr (Model::kOptimal) {
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk)
      if (live[pk]) l[pk] = lwp[pk] - model.step_opt(xpre[pk], eps[pk], ycur, xn[pk]);
  } else {
  float sc = 1.f;
  if constexpr (Model::kTimeScaled)
    if (a.step_scale) sc = a.step_scale[F.y_offset / DY + t];
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk)
    if (live[pk]) {
      if constexpr (Model::kTimeScaled)
        model.transition_s(xpre[pk], eps[pk], xn[pk], sc);
      else
        model.transition(xpre[pk], eps[pk], xn[pk]);
    }
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    if (!live[pk]) continue;
    if constexpr (GRAD) {
      float dldx, dldb;
      l[pk] = lwp[pk] - model.lik_grad(xn[pk], ycur, dldx, dldb);
      const float dxp = xpre[pk][0] + model.phi * tg[pk][0];
      const float dxs = eps[pk][0] * sc + model.phi * tg[pk][1];
      gd[pk][0] = tg[pk][2] - dldx * dxp;
      gd[pk][1] = tg[pk][3] - dldx * dxs;
      gd[pk][2] = tg[pk][4] - dldb;
      tg[pk][0] = dxp;
      tg[pk][1] = dxs;
    } else {
      l[pk] = lwp[pk] - model.lik(xn[pk], ycur);
    }
  }
  }
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk) {
    if (!live[pk]) break;
    const int64_t i = pidx(pk);
    if constexpr (GRAD) {
      tb[i * 5 + 0] = tg[pk][0];
      tb[i * 5 + 1] = tg[pk][1];
      tb[i * 5 + 2] = gd[pk][0];
      tb[i * 5 + 3] = gd[pk][1];
      tb[i * 5 + 4] = gd[pk][2];
    }
#pragma unroll
    for (int q = 0; q < D; ++q) xb[i * D + q] = xn[pk][q];
    lwb[i] = l[pk];
    if (F.out_offset >= 0) {
      const int64_t o = F.out_offset + t * n + i;
      if (a.out_value)
#pragma unroll
        for (int q = 0; q < D; ++q) a.out_value[o * D + q] = xn[pk][q];
      if (a.out_lw) a.out_lw[o] = l[pk];
      if (a.out_anc) a.out_anc[o] = (int32_t)srcs[pk];
    }
  }
#pragma unroll
  for (int pk = 0; pk < PPT; ++pk)
    if (!live[pk]) l[pk] = 0.f;  // unused lanes: defined values, masked out
  if constexpr (GRAD) {
    float tv[PPT][3];
#pragma unroll
    for (int pk = 0; pk < PPT; ++pk)
      for (int q = 0; q < 3; ++q) tv[pk][q] = live[pk] ? gd[pk][q] : 0.f;
    double g[3];
    const LseAcc r = lf_block_lse2g<PPT>(l, tv, live, s, g);
    lf_store_lse_g(a, c, r, g);
  } else {
    lf_store_lse(a, c, lf_block_lse2<PPT>(l, live, s));
  }
}

template <class Model, int D, int DY, bool GRAD>
__global__ __launch_bounds__(LT) void lf_step_kernel(LfArgs a, int64_t t, int cend_slots) {
  __shared__ LfScratch s;
  __shared__ Model model;
  __shared__ uint64_t win[kSearchWin];
  extern __shared__ uint64_t cend[];  // cend_slots entries (launch-time size)
  lf_step_chunk<Model, D, DY, GRAD>(a, t, cend_slots, blockIdx.x, s, model, win, cend);
}

// Final particles of filters whose last step wrote buffer 1 are copied to buffer 0 (as pf_kernel).
template <int D>
MOCAT_DEV void lf_finish_chunk(const LfArgs& a, int64_t c) {
  const LfDesc F = a.cdesc[c];
  if (!((F.T - 1) & 1)) return;
  for (int pk = 0; pk < PPT; ++pk) {
    const int64_t i = (c - F.chunk0) * CH + pk * LT + threadIdx.x;
    if (i >= F.n) return;
#pragma unroll
    for (int q = 0; q < D; ++q) a.x0[(F.state_offset + i) * D + q] = a.x1[(F.state_offset + i) * D + q];
    a.lw0[F.state_offset + i] = a.lw1[F.state_offset + i];
  }
}

template <int D>
__global__ __launch_bounds__(LT) void lf_finish_kernel(LfArgs a) {
  lf_finish_chunk<D>(a, blockIdx.x);
}

// Grid-wide barrier of the persistent kernel. Every workgroup adds one to the arrival counter after a
// release fence; barrier k is passed once the counter reaches k * gridDim.x, followed by an acquire
// fence. The wait is bounded: a workgroup that polls max_polls times (kBarrierPolls unless the
// MOCAT_PF_BARRIER_POLLS environment variable sets another limit, which tests use to force the
// abort) without seeing every arrival sets the abort flag, and every workgroup leaves the kernel at
// its next barrier once the flag is set (lf_abort_kernel then marks the results). All workgroups are
// resident at once (the grid is sized by the occupancy limit), so a barrier is only ever waiting on
// running workgroups.
constexpr uint32_t kBarrierPolls = 1u << 22;

MOCAT_DEV bool lf_grid_sync(uint32_t* bar, uint32_t k, uint32_t max_polls) {
  __shared__ int pass;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = k * gridDim.x;
    int ok = 1;
    for (uint32_t polls = 0;; ++polls) {
      if (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = 0;
        break;
      }
      if (polls >= max_polls) {  // max_polls = 0: every workgroup gives up at the first barrier
        __hip_atomic_store(&bar[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      if (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    pass = ok;
  }
  __syncthreads();
  return pass != 0;
}

// The phases of the persistent kernel.
template <bool GRAD>
MOCAT_DEV void lf_prep_call(const LfArgs& a, int64_t t, int64_t c, LfScratch& s) {
  lf_prep_chunk<GRAD>(a, t, c, s);
}
MOCAT_DEV void lf_csum_call(const LfArgs& a, int64_t t, int64_t c, LfScratch& s, uint64_t* cs) {
  lf_csum_chunk(a, t, c, s, cs);
}
template <class Model, int D, int DY, bool GRAD>
MOCAT_DEV void lf_step_call(const LfArgs& a, int64_t t, int cend_slots, int64_t c, LfScratch& s,
                                          Model& model, uint64_t* win, uint64_t* cend) {
  lf_step_chunk<Model, D, DY, GRAD>(a, t, cend_slots, c, s, model, win, cend);
}

// launch_mode 2: steps 1 .. max_T - 1 (after lf_init_kernel) as one persistent kernel of gridDim.x
// workgroups. Each step runs the three phases of the per-step launches (prep, prefix sums, step), a
// grid barrier after each; in a phase workgroup b takes chunks b, b + G, b + 2G, ... of the chunks
// still running. The filters are laid out in order of decreasing length, so the chunks running at
// step t are the first act[t] (act[t] = chunks of filters with T > t). The chunk functions are those
// of the per-step kernels, so the results are the same bit for bit.
template <class Model, int D, int DY, bool GRAD>
__global__ __launch_bounds__(LT) void lf_persist_kernel(LfArgs a, int64_t max_T, int cend_slots,
                                                        const int32_t* __restrict__ act, uint32_t max_polls) {
  __shared__ LfScratch s;
  __shared__ Model model;
  __shared__ uint64_t buf[kSearchWin];  // the prefix-sum phase's chunk sums, then the step's window
  extern __shared__ uint64_t cend[];
  static_assert(kSearchWin >= CH, "the LDS buffer holds a chunk of prefix sums");
  const int64_t G = gridDim.x;
  uint32_t k = 0;
  for (int64_t t = 1; t < max_T; ++t) {
    const int64_t ap = act[t - 1], as = act[t];
    for (int64_t c = blockIdx.x; c < ap; c += G) {
      __syncthreads();
      lf_prep_call<GRAD>(a, t, c, s);
    }
    if (!lf_grid_sync(a.bar, ++k, max_polls)) return;
    for (int64_t c = blockIdx.x; c < as; c += G) {
      __syncthreads();
      lf_csum_call(a, t, c, s, buf);
    }
    if (!lf_grid_sync(a.bar, ++k, max_polls)) return;
    for (int64_t c = blockIdx.x; c < as; c += G) {
      __syncthreads();
      lf_step_call<Model, D, DY, GRAD>(a, t, cend_slots, c, s, model, buf, cend);
    }
    if (!lf_grid_sync(a.bar, ++k, max_polls)) return;
  }
  const int64_t al = act[max_T - 1];
  for (int64_t c = blockIdx.x; c < al; c += G) {
    __syncthreads();
    lf_prep_call<GRAD>(a, max_T, c, s);  // the last step's statistics
  }
  if (!lf_grid_sync(a.bar, ++k, max_polls)) return;
  for (int64_t c = blockIdx.x; c < a.n_chunks; c += G) lf_finish_chunk<D>(a, c);
}

// After the persistent kernel: if a barrier gave up, every filter's log-likelihood (and gradient) is NaN.
__global__ __launch_bounds__(256) void lf_abort_kernel(LfArgs a) {
  if (__hip_atomic_load(&a.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  for (int i = threadIdx.x; i < a.n_large; i += 256) {
    const int f = a.desc[i].f;
    a.out_loglik[f] = NAN;
    if (a.out_grad)
      for (int q = 0; q < 3; ++q) a.out_grad[(int64_t)f * 3 + q] = NAN;
  }
}

// ---------------------------------------------------------------- the batched per-step hooks
// mocat's ParticleFilter.initial_sample_and_weight_vectorised / propose_and_intermediate_weight_vectorised
// (mocat/src/ssm/filtering.py:65-78, 124-139; BootstrapFilter :142-170; the optimal filter's overrides,
// mocat/src/ssm/nonlinear_gaussian.py:198-224, 264-281) and StateSpaceModel.likelihood_potential
// (mocat/src/ssm/ssm.py:73-81) for the compiled models, one thread per particle with the same model code
// and noise draws as the fused filters: particle i's noise from keys[i] (bootstrap) or row i of
// normal(keys[0], (n, D)) (optimal proposal).
//   phase 0 (initial): x[i] = initial sample; lw[i] = -likelihood(x[i], y) (0 without y; 0 for the
//                      optimal filter, which conditions its sample on y)
//   phase 1 (propose): x[i] = proposal from xprev[i]; lw[i] = intermediate log-weight (bootstrap:
//                      -likelihood(x[i], y); optimal: the predictive weight)
//   phase 2 (likelihood): lw[i] = likelihood potential of xprev[i] at y (x unused)
template <class Model, int D, int DY>
__global__ __launch_bounds__(256) void pf_hook_kernel(int phase, const float* __restrict__ params,
                                                      const float* __restrict__ xprev, const float* __restrict__ y,
                                                      const uint32_t* __restrict__ keys, float step_scale, int64_t n,
                                                      float* xout, float* lw) {
  __shared__ Model model;
  load_shared_params(model, params);
  __syncthreads();
  constexpr uint32_t hD = (D + 1) / 2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float eps[D], xp[D], x[D];
    if (phase == 2) {
#pragma unroll
      for (int q = 0; q < D; ++q) xp[q] = xprev[i * D + q];
      if constexpr (Model::kOptimal) {
        lw[i] = 0.f;  // not used: the optimal filter's likelihood hook is the bootstrap model's
      } else {
        lw[i] = model.lik(xp, y);
      }
      continue;
    }
    if constexpr (Model::kOptimal) {
      shared_eps<D>(Key{keys[0], keys[1]}, (uint32_t)(n * D), (uint32_t)i, eps);
    } else {
      const Key ki{keys[2 * i], keys[2 * i + 1]};
#pragma unroll
      for (uint32_t q = 0; q < hD; ++q) {
        uint32_t w0, w1;
        bits_pair(ki, D, q, w0, w1);
        eps[q] = normal_from_bits(w0);
        if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
      }
    }
    float l = 0.f;
    if (phase == 0) {
      if constexpr (Model::kOptimal) {
        model.init_opt(eps, y, x);
      } else {
        model.init(eps, x);
        if (y) l = -model.lik(x, y);
      }
    } else {
#pragma unroll
      for (int q = 0; q < D; ++q) xp[q] = xprev[i * D + q];
      if constexpr (Model::kOptimal) {
        l = -model.step_opt(xp, eps, y, x);
      } else {
        if constexpr (Model::kTimeScaled)
          model.transition_s(xp, eps, x, step_scale);
        else
          model.transition(xp, eps, x);
        l = -model.lik(x, y);
      }
    }
#pragma unroll
    for (int q = 0; q < D; ++q) xout[i * D + q] = x[q];
    lw[i] = l;
  }
}

template <class Model, int D, int DY>
int launch_pf_hook(int phase, const float* params, const float* xprev, const float* y, const uint32_t* keys,
                   float step_scale, int64_t n, float* x, float* lw, hipStream_t st) {
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL((pf_hook_kernel<Model, D, DY>), dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0,
                     st, phase, params, xprev, y, keys, step_scale, n, x, lw);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// Capture stream of the graph launch mode, one per device (capture needs a stream other than the
// null stream; the captured graph is then launched on the caller's stream).
hipStream_t capture_stream() {
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
    streams[dev] = nullptr;
  return streams[dev];
}

// launch_mode 0: the per-step launches go to the stream one by one, each over the chunks still
// running (act: host copy of the running-chunk counts). launch_mode 1: they are captured into one
// graph (kernel nodes only) that is instantiated and launched on the stream as a single submission;
// the call returns after the graph has run (the instantiated graph is released). When the caller's
// stream is itself being captured, the launches simply join the caller's graph. launch_mode 2: the
// init launch, then lf_persist_kernel for all steps, then lf_abort_kernel.
constexpr int kPersistBlocksPerCu = 4;

template <class Model, int D, int DY, bool GRAD = false>
int launch_pf_large(const LfArgs& a, int64_t max_T, int cend_slots, const int32_t* act_dev,
                    const std::vector<int32_t>& act, int launch_mode, hipStream_t st) {
  const dim3 chunks((unsigned)a.n_chunks), blk(LT);
  // the block-end sums are staged only by the multinomial search
  const size_t cend_bytes = a.resampling == 1 ? (size_t)cend_slots * sizeof(uint64_t) : 0;
  auto grid = [&](int64_t t) { return dim3((unsigned)std::max<int32_t>(act[t], 1)); };
  if (launch_mode == 2) {
    auto kern = lf_persist_kernel<Model, D, DY, GRAD>;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), LT,
                                                     cend_bytes) != hipSuccess)
      return MOCAT_ERR_HIP;
    if (per_cu < 1 || cus < 1) return MOCAT_ERR_SIZE;
    const int64_t G = std::min<int64_t>(a.n_chunks, (int64_t)cus * std::min(per_cu, kPersistBlocksPerCu));
    hipLaunchKernelGGL((lf_init_kernel<Model, D, DY, GRAD>), chunks, blk, 0, st, a);
    uint32_t max_polls = kBarrierPolls;
    if (const char* e = std::getenv("MOCAT_PF_BARRIER_POLLS")) max_polls = (uint32_t)std::strtoul(e, nullptr, 10);
    hipLaunchKernelGGL(kern, dim3((unsigned)G), blk, cend_bytes, st, a, max_T, cend_slots, act_dev, max_polls);
    hipLaunchKernelGGL(lf_abort_kernel, dim3(1), dim3(256), 0, st, a);
    MOCAT_LAUNCH_CHECK();
    return MOCAT_OK;
  }
  auto enqueue = [&](hipStream_t q) {
    hipLaunchKernelGGL((lf_init_kernel<Model, D, DY, GRAD>), chunks, blk, 0, q, a);
    for (int64_t t = 1; t < max_T; ++t) {
      hipLaunchKernelGGL((lf_prep_kernel<GRAD>), grid(t - 1), blk, 0, q, a, t);
      hipLaunchKernelGGL(lf_csum_kernel, grid(t), blk, 0, q, a, t);
      hipLaunchKernelGGL((lf_step_kernel<Model, D, DY, GRAD>), grid(t), blk, cend_bytes, q, a, t, cend_slots);
    }
    hipLaunchKernelGGL((lf_prep_kernel<GRAD>), grid(max_T - 1), blk, 0, q, a, max_T);  // the last step's statistics
    hipLaunchKernelGGL((lf_finish_kernel<D>), chunks, blk, 0, q, a);
  };
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (launch_mode == 1 && hipStreamIsCapturing(st, &cap) != hipSuccess) return MOCAT_ERR_HIP;
  if (launch_mode == 0 || cap != hipStreamCaptureStatusNone) {
    enqueue(st);
    MOCAT_LAUNCH_CHECK();
    return MOCAT_OK;
  }
  hipStream_t cs = capture_stream();
  if (!cs) return MOCAT_ERR_HIP;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return MOCAT_ERR_HIP;
  enqueue(cs);
  const hipError_t le = hipGetLastError();
  if (hipStreamEndCapture(cs, &graph) != hipSuccess || le != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    return MOCAT_ERR_HIP;
  }
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return MOCAT_ERR_HIP;
  e = hipGraphLaunch(exec, st);
  const hipError_t se = hipStreamSynchronize(st);
  (void)hipGraphExecDestroy(exec);
  return (e == hipSuccess && se == hipSuccess) ? MOCAT_OK : MOCAT_ERR_HIP;
}

template <class Model, int D, int DY, bool GRAD = false>
int launch_pf(const PfArgs& a, hipStream_t st) {
  const int grid = a.n_filters < 4096 ? a.n_filters : 4096;
  hipLaunchKernelGGL((pf_kernel<Model, D, DY, GRAD>), dim3(grid), dim3(NT), 0, st, a);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}


struct LfLayout {
  int n_large = 0;
  int cend_slots = 1;  // SG-block end sums the step kernel stages: the largest filter's, up to kChunkEnds
  int64_t n_chunks = 0, max_T = 0;
  int64_t n_tk = 0;  // step-key slots: the sum of T over the flagged filters
  int64_t o_desc, o_cf, o_m, o_s1, o_s2, o_g, o_tot, o_st, o_cend, o_bnd, o_keys, o_act, o_bar, o_pk, bytes;
};

int64_t lf_align(int64_t v) { return (v + 255) / 256 * 256; }

LfLayout lf_layout(const mocat_pf_filter* filters, int n_filters) {
  LfLayout L;
  for (int f = 0; f < n_filters; ++f) {
    if (!(filters[f].flags & kPfLarge)) continue;
    ++L.n_large;
    L.n_chunks += (filters[f].n + CH - 1) / CH;
    L.max_T = std::max<int64_t>(L.max_T, filters[f].T);
    L.n_tk += std::max<int64_t>(filters[f].T, 0);
    L.cend_slots = (int)std::min<int64_t>(kChunkEnds, std::max<int64_t>(L.cend_slots, (filters[f].n + SG - 1) / SG));
  }
  int64_t o = 0;
  L.o_desc = o;
  o = lf_align(o + (int64_t)L.n_large * (int64_t)sizeof(LfDesc));
  L.o_cf = o;
  o = lf_align(o + L.n_chunks * (int64_t)sizeof(LfDesc));
  L.o_m = o;
  o = lf_align(o + L.n_chunks * 4);
  L.o_s1 = o;
  o = lf_align(o + L.n_chunks * 8);
  L.o_s2 = o;
  o = lf_align(o + L.n_chunks * 8);
  L.o_g = o;
  o = lf_align(o + L.n_chunks * 24);
  L.o_tot = o;
  o = lf_align(o + L.n_chunks * 8);
  L.o_st = o;
  o = lf_align(o + (int64_t)L.n_large * (int64_t)sizeof(LfState));
  L.o_cend = o;
  o = lf_align(o + L.n_chunks * (CH / SG) * 8);
  L.o_bnd = o;
  o = lf_align(o + L.n_chunks * 8);
  L.o_keys = o;
  o = lf_align(o + L.n_tk * (int64_t)sizeof(StepKey));
  L.o_act = o;
  o = lf_align(o + (L.max_T + 1) * 4);
  L.o_bar = o;
  o = lf_align(o + 8);
  L.o_pk = o;
  o = lf_align(o + L.n_chunks * CH * 8);
  L.bytes = o;
  return L;
}

}  // namespace

namespace {

// The two launches of resampling from weights: exact prefix sums into ws (n uint64, every entry
// written before it is read), then the ancestor search.
int launch_resample(const float* weights, int64_t n, int method, int from_log, const uint32_t* key, uint64_t* ws,
                    int32_t* idx, hipStream_t stream) {
  hipLaunchKernelGGL(resample_csum_kernel, dim3(1), dim3(NT), 0, stream, weights, n, from_log, ws);
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(resample_search_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, stream,
                     (const uint64_t*)ws, n, key, method, idx);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

// Scratch memory of mocat_resample_noalloc: the prefix sums of up to kScratchMaxWeights weights (8 MiB),
// one copy per GPU, held until the process ends. Host code never passes the array's name to a kernel
// (in host code it is a stand-in, not a device address): each device's copy is addressed through
// the pointer hipGetSymbolAddress returns with that device current, looked up once per device.
constexpr int64_t kScratchMaxWeights = int64_t(1) << 20;
static __device__ __align__(256) uint64_t scratch[kScratchMaxWeights];
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mutex;
uint64_t* g_scratch_ptr[kScratchMaxDevices] = {};

int scratch_on_current_device(uint64_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchMaxDevices) return MOCAT_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MOCAT_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MOCAT_OK;
}

}  // namespace

// (d, dy) pairs compiled for the linear-Gaussian models
#define MOCAT_PF_LG_SHAPES                                                                                  \
  MOCAT_PF_CASE(1, 1) MOCAT_PF_CASE(2, 2) MOCAT_PF_CASE(4, 4) MOCAT_PF_CASE(4, 2) MOCAT_PF_CASE(5, 5) \
  MOCAT_PF_CASE(8, 8) MOCAT_PF_CASE(2, 1) MOCAT_PF_CASE(3, 3) MOCAT_PF_CASE(4, 1) MOCAT_PF_CASE(8, 4)

extern "C" {

int mocat_pf_param_count(int model, int d, int dy, int64_t* count) {
  if (!count) return MOCAT_ERR_ARG;
  if (model == 1) {
    if (d != dy || d < 1 || d > 4) return MOCAT_ERR_ARG;
    *count = d == 1 ? Sv::NPARAM : 2 * d + 3 * d * d;
    return MOCAT_OK;
  }
  // models 0 and 2: the compiled shapes, or the runtime-dimension hooks up to 16
  if ((model != 0 && model != 2) || d < 1 || d > 16 || dy < 1 || dy > 16) return MOCAT_ERR_ARG;
  *count = model == 0 ? d + 3 * d * d + dy * d + dy * dy + 1 : d + 3 * dy * d + 3 * d * d + dy * dy + 1;
  return MOCAT_OK;
}

int mocat_pf_compiled_shape(int model, int d, int dy, int* compiled) {
  if (!compiled) return MOCAT_ERR_ARG;
  int c = 0;
  if (model == 1) c = d == dy && d >= 1 && d <= 4;
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) c = 1;
  if (model == 0 || model == 2) {
    MOCAT_PF_LG_SHAPES
  }
#undef MOCAT_PF_CASE
  *compiled = c;
  return MOCAT_OK;
}

int mocat_pf_run(int model, int d, int dy, const mocat_pf_filter* host_filters, const mocat_pf_filter* filters,
                 int n_filters, const float* params,
                 const float* y, const float* step_scale, float* x_ws0, float* x_ws1, float* lw_ws0, float* lw_ws1, uint64_t* csum_ws,
                 int32_t* anc_ws, int resampling, float ess_threshold, float* out_value, float* out_lw,
                 int32_t* out_anc, float* out_ess, double* out_loglik_t, double* out_loglik, float* tan_ws0,
                 float* tan_ws1, double* out_grad, hipStream_t stream) {
  if (n_filters < 0 || !host_filters || !filters || !params || !y || !x_ws0 || !x_ws1 || !lw_ws0 || !lw_ws1 ||
      !csum_ws || !anc_ws || !out_loglik || resampling < 0 || resampling > 2)
    return MOCAT_ERR_ARG;
  if (out_grad && (model != 1 || !tan_ws0 || !tan_ws1)) return MOCAT_ERR_ARG;
  if (n_filters == 0) return MOCAT_OK;
  // argument checks read the host copy only; the kernel reads the device copy
  for (int f = 0; f < n_filters; ++f) {
    const mocat_pf_filter& F = host_filters[f];
    if (F.flags & kPfLarge) continue;
    if (F.n <= 0 || F.T <= 0 || F.n > 0x7FFFFFFFll || F.T > 0xFFFFFFFFll || F.state_offset < 0 ||
        F.y_offset < 0 || F.param_offset < 0)
      return MOCAT_ERR_ARG;
    if ((out_ess || out_loglik_t) && F.t_offset < 0) return MOCAT_ERR_ARG;
    if (out_grad && (F.flags & kPfResume)) return MOCAT_ERR_ARG;
    if (resampling == 2 && F.n * F.n >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
    if (model == 2 && F.n * (int64_t)d > 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  }
  PfArgs a{filters, n_filters, params, y, step_scale, x_ws0, x_ws1, lw_ws0, lw_ws1, csum_ws, anc_ws, resampling,
           ess_threshold, out_value, out_lw, out_anc, out_ess, out_loglik_t, out_loglik, tan_ws0, tan_ws1, out_grad};
  if (model == 1) {
    if (d != dy) return MOCAT_ERR_ARG;
    if (d == 1) return out_grad ? launch_pf<Sv, 1, 1, true>(a, stream) : launch_pf<Sv, 1, 1>(a, stream);
    if (out_grad) return MOCAT_ERR_ARG;  // the gradient is compiled for the 1-D model
    if (d == 2) return launch_pf<SvD<2>, 2, 2>(a, stream);
    if (d == 3) return launch_pf<SvD<3>, 3, 3>(a, stream);
    if (d == 4) return launch_pf<SvD<4>, 4, 4>(a, stream);
    return MOCAT_ERR_ARG;
  }
  if (model == 2) {
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf<LgssmOpt<DD, YY>, DD, YY>(a, stream);
    MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
    return MOCAT_ERR_ARG;
  }
  if (model != 0) return MOCAT_ERR_ARG;
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf<Lgssm<DD, YY>, DD, YY>(a, stream);
  MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
  return MOCAT_ERR_ARG;
}

int mocat_pf_large_workspace_bytes(const mocat_pf_filter* host_filters, int n_filters, int64_t* bytes) {
  if (!host_filters || n_filters < 0 || !bytes) return MOCAT_ERR_ARG;
  *bytes = lf_layout(host_filters, n_filters).bytes;
  return MOCAT_OK;
}

int mocat_pf_run_large(int model, int d, int dy, const mocat_pf_filter* host_filters, int n_filters,
                       const float* params, const float* y, const float* step_scale, float* x_ws0, float* x_ws1, float* lw_ws0, float* lw_ws1, uint64_t* csum_ws,
                       int resampling, float ess_threshold, float* out_value, float* out_lw, int32_t* out_anc,
                       float* out_ess, double* out_loglik_t, double* out_loglik, float* tan_ws0, float* tan_ws1,
                       double* out_grad, void* workspace, int64_t workspace_bytes, int launch_mode,
                       hipStream_t stream) {
  if (n_filters < 0 || !host_filters || !params || !y || !x_ws0 || !x_ws1 || !lw_ws0 || !lw_ws1 || !csum_ws ||
      !out_loglik || resampling < 0 || resampling > 2 || launch_mode < 0 || launch_mode > 2)
    return MOCAT_ERR_ARG;
  if (out_grad && (model != 1 || !tan_ws0 || !tan_ws1)) return MOCAT_ERR_ARG;
  const LfLayout L = lf_layout(host_filters, n_filters);
  if (L.n_large == 0) return MOCAT_OK;
  if (!workspace || workspace_bytes < L.bytes) return MOCAT_ERR_SIZE;
  if (L.n_chunks > 0x7FFFFFFFll || L.n_large > 0x7FFFFFFF) return MOCAT_ERR_SIZE;
  std::vector<LfDesc> desc;
  std::vector<LfDesc> cdesc;
  desc.reserve(L.n_large);
  cdesc.reserve(L.n_chunks);
  // filters in order of decreasing length (ties in input order): the chunks still running at any step
  // are then a prefix of the chunk list
  std::vector<int> order;
  for (int f = 0; f < n_filters; ++f)
    if (host_filters[f].flags & kPfLarge) order.push_back(f);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return host_filters[x].T > host_filters[y].T; });
  int64_t chunk0 = 0, key_slot = 0;
  for (const int f : order) {
    const mocat_pf_filter& F = host_filters[f];
    if (F.n <= 0 || F.T <= 0 || F.n > 0x7FFFFFFFll || F.T > 0xFFFFFFFFll || F.state_offset < 0 ||
        F.y_offset < 0 || F.param_offset < 0)
      return MOCAT_ERR_ARG;
    if ((out_ess || out_loglik_t) && F.t_offset < 0) return MOCAT_ERR_ARG;
    if (resampling == 2 && F.n * F.n >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
    if (model == 2 && F.n * (int64_t)d > 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
    if (out_grad && (F.flags & kPfResume)) return MOCAT_ERR_ARG;
    const int64_t nchunk = (F.n + CH - 1) / CH;
    desc.push_back(LfDesc{F.n, F.T, F.y_offset, F.state_offset, F.out_offset, F.t_offset, chunk0, key_slot,
                          {F.key[0], F.key[1]}, F.param_offset, f, F.flags, (int32_t)desc.size()});
    key_slot += F.T;
    for (int64_t k = 0; k < nchunk; ++k) cdesc.push_back(desc.back());
    chunk0 += nchunk;
  }
  // act[t]: chunks of the filters with T > t, t = 0 .. max_T
  std::vector<int32_t> act((size_t)L.max_T + 1, 0);
  for (const LfDesc& F : desc) {
    const int32_t nchunk = (int32_t)((F.n + CH - 1) / CH);
    for (int64_t t = 0; t < F.T; ++t) act[(size_t)t] += nchunk;
  }
  char* ws = static_cast<char*>(workspace);
  if (hipMemcpyAsync(ws + L.o_desc, desc.data(), desc.size() * sizeof(LfDesc), hipMemcpyHostToDevice, stream) !=
          hipSuccess ||
      hipMemcpyAsync(ws + L.o_cf, cdesc.data(), cdesc.size() * sizeof(LfDesc), hipMemcpyHostToDevice,
                     stream) != hipSuccess ||
      hipMemcpyAsync(ws + L.o_act, act.data(), act.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream) !=
          hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)  // the host vectors are released on return
    return MOCAT_ERR_HIP;
  const int32_t* act_dev = reinterpret_cast<const int32_t*>(ws + L.o_act);
  LfArgs a{reinterpret_cast<const LfDesc*>(ws + L.o_desc), reinterpret_cast<const LfDesc*>(ws + L.o_cf), L.n_large,
           L.n_chunks, params, y, step_scale, x_ws0, x_ws1, lw_ws0, lw_ws1, csum_ws, tan_ws0, tan_ws1, resampling,
           ess_threshold, out_value, out_lw, out_anc, out_ess, out_loglik_t, out_loglik, out_grad,
           reinterpret_cast<float*>(ws + L.o_m), reinterpret_cast<double*>(ws + L.o_s1),
           reinterpret_cast<double*>(ws + L.o_s2), reinterpret_cast<double*>(ws + L.o_g),
           reinterpret_cast<uint64_t*>(ws + L.o_tot), reinterpret_cast<LfState*>(ws + L.o_st),
           reinterpret_cast<uint64_t*>(ws + L.o_cend), reinterpret_cast<int64_t*>(ws + L.o_bnd),
           reinterpret_cast<StepKey*>(ws + L.o_keys), reinterpret_cast<uint32_t*>(ws + L.o_bar),
           model == 2 ? nullptr : reinterpret_cast<uint32_t*>(ws + L.o_pk)};
  if (model == 1) {
    if (d != dy) return MOCAT_ERR_ARG;
    if (d != 1 && out_grad) return MOCAT_ERR_ARG;
    if (d == 2) return launch_pf_large<SvD<2>, 2, 2>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
    if (d == 3) return launch_pf_large<SvD<3>, 3, 3>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
    if (d == 4) return launch_pf_large<SvD<4>, 4, 4>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
    if (d != 1) return MOCAT_ERR_ARG;
    return out_grad ? launch_pf_large<Sv, 1, 1, true>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream)
                    : launch_pf_large<Sv, 1, 1>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
  }
  if (model == 2) {
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf_large<LgssmOpt<DD, YY>, DD, YY>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
    MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
    return MOCAT_ERR_ARG;
  }
  if (model != 0) return MOCAT_ERR_ARG;
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf_large<Lgssm<DD, YY>, DD, YY>(a, L.max_T, L.cend_slots, act_dev, act, launch_mode, stream);
  MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
  return MOCAT_ERR_ARG;
}

int mocat_pf_hook(int model, int d, int dy, int phase, const float* params, const float* x_prev, const float* y,
                  const uint32_t* keys, float step_scale, int64_t n, float* x, float* log_weight,
                  hipStream_t stream) {
  if (!params || !log_weight || n < 0 || phase < 0 || phase > 2 || !(step_scale >= 0.f)) return MOCAT_ERR_ARG;
  if (n == 0) return MOCAT_OK;
  if (phase != 2 && (!keys || !x)) return MOCAT_ERR_ARG;
  if (phase != 0 && !x_prev) return MOCAT_ERR_ARG;
  if (phase != 0 && !y) return MOCAT_ERR_ARG;
  if (model == 2 && (phase == 2 || !y)) return MOCAT_ERR_ARG;
  if (n > 0x7FFFFFFFll || (model == 2 && n * (int64_t)d > 0xFFFFFFFFll)) return MOCAT_ERR_SIZE;
  if (model == 1) {
    if (d != dy) return MOCAT_ERR_ARG;
    if (d == 1) return launch_pf_hook<Sv, 1, 1>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
    if (d == 2) return launch_pf_hook<SvD<2>, 2, 2>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
    if (d == 3) return launch_pf_hook<SvD<3>, 3, 3>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
    if (d == 4) return launch_pf_hook<SvD<4>, 4, 4>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
    return MOCAT_ERR_ARG;
  }
  if (model == 2) {
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf_hook<LgssmOpt<DD, YY>, DD, YY>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
    MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
    return launch_lg_generic_hook(2, phase, d, dy, params, x_prev, y, keys, n, x, log_weight, stream);
  }
  if (model != 0) return MOCAT_ERR_ARG;
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf_hook<Lgssm<DD, YY>, DD, YY>(phase, params, x_prev, y, keys, step_scale, n, x, log_weight, stream);
  MOCAT_PF_LG_SHAPES
#undef MOCAT_PF_CASE
  // another (d, d_y): the runtime-dimension kernel (d, d_y up to 16)
  return launch_lg_generic_hook(0, phase, d, dy, params, x_prev, y, keys, n, x, log_weight, stream);
}

int mocat_resample(const float* weights, int64_t n, int method, int from_log, const uint32_t* key, void* workspace,
                   int32_t* idx, hipStream_t stream) {
  if (!weights || !key || !workspace || !idx || n <= 0 || n > 0x7FFFFFFFll || method < 0 || method > 1)
    return MOCAT_ERR_ARG;
  return launch_resample(weights, n, method, from_log, key, static_cast<uint64_t*>(workspace), idx, stream);
}

int mocat_resample_noalloc(const float* weights, int64_t n, int method, int from_log, const uint32_t* key,
                           int32_t* idx, hipStream_t stream) {
  if (!weights || !key || !idx || n <= 0 || method < 0 || method > 1) return MOCAT_ERR_ARG;
  if (n > kScratchMaxWeights) return MOCAT_ERR_SIZE;
  uint64_t* ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc != MOCAT_OK) return rc;
  return launch_resample(weights, n, method, from_log, key, ws, idx, stream);
}

int mocat_resample_categorical(const float* log_weight, int64_t n, int64_t n_out, const uint32_t* key, int32_t* idx,
                               hipStream_t stream) {
  if (!log_weight || !key || !idx || n <= 0 || n_out <= 0) return MOCAT_ERR_ARG;
  if (n * n_out >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  const int64_t blocks = (n_out + 3) / 4;  // one wave per output, four per workgroup
  hipLaunchKernelGGL(categorical_kernel, dim3((unsigned)(blocks < 65535 ? blocks : 65535)), dim3(256), 0, stream,
                     log_weight, n, n_out, key, idx);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
