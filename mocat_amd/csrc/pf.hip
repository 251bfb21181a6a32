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
#include "common.hpp"

using namespace mocat;

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
  int32_t reserved;
} mocat_pf_filter;
}

namespace {

constexpr int NT = 512;
constexpr int NW = NT / kWave;
constexpr double kLog2PiD = 1.83787706640934548356;

// ---------------------------------------------------------------- models
// Linear Gaussian (mocat/src/ssm/linear_gaussian/linear_gaussian.py:46-51, 86-94, 118-128):
// params: m0[D], L0[D*D], A[D*D], LQ[D*D], H[DY*D], SR[DY*DY], nlz
//   x0 = L0 @ eps + m0 ; x' = A @ x + LQ @ eps ; lik = 0.5 |(y - H x) @ SR|^2 + nlz
template <int D, int DY>
struct Lgssm {
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
    }
    float s = 0.f;
    #pragma unroll
    for (int k = 0; k < DY; ++k) {
      float v = 0.f;
      #pragma unroll
      for (int j = 0; j < DY; ++j) v += diff[j] * SR[j * DY + k];
      s += v * v;
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

// ---------------------------------------------------------------- block helpers
struct LseAcc {
  float m;
  double s1, s2;
};

MOCAT_DEV LseAcc lse_combine(LseAcc a, LseAcc b) {
  LseAcc r;
  r.m = fmaxf(a.m, b.m);
  if (r.m == -INFINITY) {
    r.s1 = 0.0;
    r.s2 = 0.0;
    return r;
  }
  const double fa = a.m == -INFINITY ? 0.0 : exp((double)a.m - (double)r.m);
  const double fb = b.m == -INFINITY ? 0.0 : exp((double)b.m - (double)r.m);
  r.s1 = a.s1 * fa + b.s1 * fb;
  r.s2 = a.s2 * fa * fa + b.s2 * fb * fb;
  return r;
}

MOCAT_DEV void lse_add(LseAcc& a, float v) {
  if (v == -INFINITY || isnan(v)) {
    if (isnan(v)) a.m = NAN;
    return;
  }
  if (v > a.m) {
    const double f = a.m == -INFINITY ? 0.0 : exp((double)a.m - (double)v);
    a.s1 = a.s1 * f + 1.0;
    a.s2 = a.s2 * f * f + 1.0;
    a.m = v;
  } else {
    const double e = exp((double)v - (double)a.m);
    a.s1 += e;
    a.s2 += e * e;
  }
}

struct GradAcc {
  float m;
  double s1, g0, g1, g2;
};

MOCAT_DEV GradAcc grad_combine(GradAcc a, GradAcc b) {
  GradAcc r;
  r.m = fmaxf(a.m, b.m);
  if (r.m == -INFINITY) return GradAcc{-INFINITY, 0.0, 0.0, 0.0, 0.0};
  const double fa = a.m == -INFINITY ? 0.0 : exp((double)a.m - (double)r.m);
  const double fb = b.m == -INFINITY ? 0.0 : exp((double)b.m - (double)r.m);
  r.s1 = a.s1 * fa + b.s1 * fb;
  r.g0 = a.g0 * fa + b.g0 * fb;
  r.g1 = a.g1 * fa + b.g1 * fb;
  r.g2 = a.g2 * fa + b.g2 * fb;
  return r;
}

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
MOCAT_DEV void block_lse(LseAcc a, BlockScratch& s, double& lse, double& lse2) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int o = kWave / 2; o > 0; o >>= 1) {
    LseAcc b;
    b.m = __shfl_xor(a.m, o, kWave);
    b.s1 = __shfl_xor(a.s1, o, kWave);
    b.s2 = __shfl_xor(a.s2, o, kWave);
    a = lse_combine(a, b);
  }
  __syncthreads();
  if (lane == 0) {
    s.fm[w] = a.m;
    s.d1[w] = a.s1;
    s.d2[w] = a.s2;
  }
  __syncthreads();
  LseAcc r{s.fm[0], s.d1[0], s.d2[0]};
  for (int i = 1; i < NW; ++i) r = lse_combine(r, LseAcc{s.fm[i], s.d1[i], s.d2[i]});
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
}

// Weighted means sum_i w_i t_i / sum_i w_i of the three tangents, w_i = exp(lw_i - max), over the block.
MOCAT_DEV void block_grad(GradAcc a, BlockScratch& s, double out[3]) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int o = kWave / 2; o > 0; o >>= 1) {
    GradAcc b;
    b.m = __shfl_xor(a.m, o, kWave);
    b.s1 = __shfl_xor(a.s1, o, kWave);
    b.g0 = __shfl_xor(a.g0, o, kWave);
    b.g1 = __shfl_xor(a.g1, o, kWave);
    b.g2 = __shfl_xor(a.g2, o, kWave);
    a = grad_combine(a, b);
  }
  __syncthreads();
  if (lane == 0) {
    s.fm[w] = a.m;
    s.d1[w] = a.s1;
    s.g0[w] = a.g0;
    s.g1[w] = a.g1;
    s.g2[w] = a.g2;
  }
  __syncthreads();
  GradAcc r{s.fm[0], s.d1[0], s.g0[0], s.g1[0], s.g2[0]};
  for (int i = 1; i < NW; ++i) r = grad_combine(r, GradAcc{s.fm[i], s.d1[i], s.g0[i], s.g1[i], s.g2[i]});
  out[0] = r.g0 / r.s1;
  out[1] = r.g1 / r.s1;
  out[2] = r.g2 / r.s1;
}

MOCAT_DEV void grad_add(GradAcc& a, float v, float t0, float t1, float t2) {
  if (!(v > -INFINITY)) return;
  if (v > a.m) {
    const double f = a.m == -INFINITY ? 0.0 : exp((double)a.m - (double)v);
    a.s1 = a.s1 * f + 1.0;
    a.g0 = a.g0 * f + t0;
    a.g1 = a.g1 * f + t1;
    a.g2 = a.g2 * f + t2;
    a.m = v;
  } else {
    const double e = exp((double)v - (double)a.m);
    a.s1 += e;
    a.g0 += e * t0;
    a.g1 += e * t1;
    a.g2 += e * t2;
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

// Ancestor indices for weights w_i = weight(i) (all threads participate).
// method 0 systematic, 1 multinomial (uniforms from key). csum: n uint64 of workspace.
template <class WeightFn>
MOCAT_DEV void resample_block(WeightFn weight, int64_t n, Key key, int method, uint64_t* __restrict__ csum,
                              int32_t* __restrict__ anc, BlockScratch& s) {
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
  __threadfence_block();
  __syncthreads();
  const double W = (double)total;
  if (method == 0) {
    const double u = (double)uniform_from_bits(bits_at(key, 1, 0), 0.f, 1.f);
    const double step = W / (double)n;
    for (int64_t j = t; j < n; j += NT) anc[j] = (int32_t)search_csum(csum, n, ((double)j + u) * step);
  } else {
    const uint32_t S = (uint32_t)n;
    for (int64_t j = t; j < n; j += NT) {
      const double u = (double)uniform_from_bits(bits_at(key, S, (uint32_t)j), 0.f, 1.f);
      anc[j] = (int32_t)search_csum(csum, n, u * W);
    }
  }
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
    const int64_t n = F.n, T = F.T;
    __syncthreads();
    if (tid == 0) model.load(a.params + F.param_offset);
    __syncthreads();
    float* xa = a.x0 + F.state_offset * D;
    float* xb = a.x1 + F.state_offset * D;
    float* lwa = a.lw0 + F.state_offset;
    float* lwb = a.lw1 + F.state_offset;
    uint64_t* csum = a.csum + F.state_offset;
    int32_t* anc = a.anc + F.state_offset;
    const float* yf = a.y + F.y_offset;
    const Key key{F.key[0], F.key[1]};
    const Key rkey = split_at(key, 2, 0), sub = split_at(key, 2, 1);
    constexpr uint32_t hD = (D + 1) / 2;
    float* ta = GRAD ? a.tan0 + F.state_offset * 5 : nullptr;
    float* tb = GRAD ? a.tan1 + F.state_offset * 5 : nullptr;
    GradAcc gacc{-INFINITY, 0.0, 0.0, 0.0, 0.0};
    double gtot[3] = {0.0, 0.0, 0.0}, gprev[3] = {0.0, 0.0, 0.0};

    // ---- t = 0: initiate_particles
    LseAcc acc{-INFINITY, 0.0, 0.0};
    for (int64_t i = tid; i < n; i += NT) {
      const Key ki = split_at(sub, (uint32_t)n, (uint32_t)i);
      float eps[D], x[D];
      #pragma unroll
      for (uint32_t q = 0; q < hD; ++q) {
        uint32_t w0, w1;
        bits_pair(ki, D, q, w0, w1);
        eps[q] = normal_from_bits(w0);
        if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
      }
      model.init(eps, x);
      float l;
      if constexpr (GRAD) {
        float dldx, dldb;
        l = -model.lik_grad(x, yf, dldx, dldb);
        // the initial state does not depend on (phi, sigma)
        ta[i * 5 + 0] = 0.f;
        ta[i * 5 + 1] = 0.f;
        ta[i * 5 + 2] = 0.f;
        ta[i * 5 + 3] = 0.f;
        ta[i * 5 + 4] = -dldb;
        grad_add(gacc, l, 0.f, 0.f, -dldb);
      } else {
        l = -model.lik(x, yf);
      }
      #pragma unroll
      for (int q = 0; q < D; ++q) xa[i * D + q] = x[q];
      lwa[i] = l;
      lse_add(acc, l);
      if (a.out_value && F.out_offset >= 0)
        #pragma unroll
        for (int q = 0; q < D; ++q) a.out_value[(F.out_offset + i) * D + q] = x[q];
      if (a.out_lw && F.out_offset >= 0) a.out_lw[F.out_offset + i] = l;
      if (a.out_anc && F.out_offset >= 0) a.out_anc[F.out_offset + i] = (int32_t)i;
    }
    double lse, lse2;
    block_lse(acc, s, lse, lse2);
    if constexpr (GRAD) {
      block_grad(gacc, s, gprev);
      for (int k = 0; k < 3; ++k) gtot[k] = gprev[k];
    }
    const double logn = log((double)n);
    double loglik = lse - logn;
    double ess = exp(2.0 * lse - lse2);
    if (tid == 0) {
      if (a.out_ess) a.out_ess[F.t_offset] = (float)ess;
      if (a.out_loglik_t) a.out_loglik_t[F.t_offset] = loglik;
    }

    // ---- t = 1 .. T-1
    for (int64_t t = 1; t < T; ++t) {
      const Key ik = split_at(rkey, (uint32_t)(T - 1), (uint32_t)(t - 1));
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
      acc = LseAcc{-INFINITY, 0.0, 0.0};
      gacc = GradAcc{-INFINITY, 0.0, 0.0, 0.0, 0.0};
      for (int64_t i = tid; i < n; i += NT) {
        const int64_t src = res ? (int64_t)anc[i] : i;
        const Key ki = split_at(ik, (uint32_t)n, (uint32_t)i);
        float eps[D], xp[D], x[D];
        #pragma unroll
        for (uint32_t q = 0; q < hD; ++q) {
          uint32_t w0, w1;
          bits_pair(ki, D, q, w0, w1);
          eps[q] = normal_from_bits(w0);
          if (q + hD < (uint32_t)D) eps[q + hD] = normal_from_bits(w1);
        }
        #pragma unroll
        for (int q = 0; q < D; ++q) xp[q] = xa[src * D + q];
        model.transition(xp, eps, x);
        float l;
        if constexpr (GRAD) {
          float dldx, dldb;
          l = (res ? 0.f : lwa[i]) - model.lik_grad(x, ycur, dldx, dldb);
          const float dxp = xp[0] + model.phi * ta[src * 5 + 0];
          const float dxs = eps[0] + model.phi * ta[src * 5 + 1];
          const float d0 = (res ? 0.f : ta[i * 5 + 2]) - dldx * dxp;
          const float d1 = (res ? 0.f : ta[i * 5 + 3]) - dldx * dxs;
          const float d2 = (res ? 0.f : ta[i * 5 + 4]) - dldb;
          tb[i * 5 + 0] = dxp;
          tb[i * 5 + 1] = dxs;
          tb[i * 5 + 2] = d0;
          tb[i * 5 + 3] = d1;
          tb[i * 5 + 4] = d2;
          grad_add(gacc, l, d0, d1, d2);
        } else {
          l = (res ? 0.f : lwa[i]) - model.lik(x, ycur);
        }
        #pragma unroll
        for (int q = 0; q < D; ++q) xb[i * D + q] = x[q];
        lwb[i] = l;
        lse_add(acc, l);
        if (F.out_offset >= 0) {
          const int64_t o = F.out_offset + t * n + i;
          if (a.out_value)
            #pragma unroll
            for (int q = 0; q < D; ++q) a.out_value[o * D + q] = x[q];
          if (a.out_lw) a.out_lw[o] = l;
          if (a.out_anc) a.out_anc[o] = (int32_t)src;
        }
      }
      block_lse(acc, s, lse, lse2);
      loglik += lse - lse_res;
      if constexpr (GRAD) {
        double gnow[3];
        block_grad(gacc, s, gnow);
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
// from_log: w holds log-weights, the weights used are exp(w - max w) (f32).
__global__ __launch_bounds__(NT) void resample_kernel(const float* __restrict__ w, int64_t n, const uint32_t* key,
                                                      int method, int from_log, uint64_t* csum, int32_t* idx) {
  __shared__ BlockScratch s;
  const Key k{key[0], key[1]};
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
  resample_block(wfn, n, k, method, csum, idx, s);
}

__global__ __launch_bounds__(256) void categorical_kernel(const float* __restrict__ lw, int64_t n, int64_t n_out,
                                                          const uint32_t* key, int32_t* idx) {
  const Key k{key[0], key[1]};
  const uint32_t S = (uint32_t)(n * n_out);
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n_out; j += (int64_t)gridDim.x * blockDim.x)
    idx[j] = categorical_row(lw, n, k, S, j);
}

template <class Model, int D, int DY, bool GRAD = false>
int launch_pf(const PfArgs& a, hipStream_t st) {
  const int grid = a.n_filters < 4096 ? a.n_filters : 4096;
  hipLaunchKernelGGL((pf_kernel<Model, D, DY, GRAD>), dim3(grid), dim3(NT), 0, st, a);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // namespace

extern "C" {

int mocat_pf_param_count(int model, int d, int dy, int64_t* count) {
  if (!count) return MOCAT_ERR_ARG;
  if (model == 1) {
    *count = Sv::NPARAM;
    return MOCAT_OK;
  }
  if (model != 0 || d < 1 || d > 8 || dy < 1 || dy > 8) return MOCAT_ERR_ARG;
  *count = d + 3 * d * d + dy * d + dy * dy + 1;
  return MOCAT_OK;
}

int mocat_pf_run(int model, int d, int dy, const mocat_pf_filter* filters, int n_filters, const float* params,
                 const float* y, float* x_ws0, float* x_ws1, float* lw_ws0, float* lw_ws1, uint64_t* csum_ws,
                 int32_t* anc_ws, int resampling, float ess_threshold, float* out_value, float* out_lw,
                 int32_t* out_anc, float* out_ess, double* out_loglik_t, double* out_loglik, float* tan_ws0,
                 float* tan_ws1, double* out_grad, hipStream_t stream) {
  if (n_filters < 0 || !filters || !params || !y || !x_ws0 || !x_ws1 || !lw_ws0 || !lw_ws1 || !csum_ws || !anc_ws ||
      !out_loglik || resampling < 0 || resampling > 2)
    return MOCAT_ERR_ARG;
  if (out_grad && (model != 1 || !tan_ws0 || !tan_ws1)) return MOCAT_ERR_ARG;
  if (n_filters == 0) return MOCAT_OK;
  PfArgs a{filters, n_filters, params, y, x_ws0, x_ws1, lw_ws0, lw_ws1, csum_ws, anc_ws, resampling,
           ess_threshold, out_value, out_lw, out_anc, out_ess, out_loglik_t, out_loglik, tan_ws0, tan_ws1, out_grad};
  if (model == 1) {
    if (d != 1 || dy != 1) return MOCAT_ERR_ARG;
    return out_grad ? launch_pf<Sv, 1, 1, true>(a, stream) : launch_pf<Sv, 1, 1>(a, stream);
  }
  if (model != 0) return MOCAT_ERR_ARG;
#define MOCAT_PF_CASE(DD, YY) \
  if (d == DD && dy == YY) return launch_pf<Lgssm<DD, YY>, DD, YY>(a, stream);
  MOCAT_PF_CASE(1, 1) MOCAT_PF_CASE(2, 2) MOCAT_PF_CASE(4, 4) MOCAT_PF_CASE(4, 2) MOCAT_PF_CASE(5, 5)
  MOCAT_PF_CASE(8, 8) MOCAT_PF_CASE(2, 1) MOCAT_PF_CASE(3, 3) MOCAT_PF_CASE(4, 1) MOCAT_PF_CASE(8, 4)
#undef MOCAT_PF_CASE
  return MOCAT_ERR_ARG;
}

int mocat_resample(const float* weights, int64_t n, int method, int from_log, const uint32_t* key, void* workspace,
                   int32_t* idx, hipStream_t stream) {
  if (!weights || !key || !workspace || !idx || n <= 0 || n > 0x7FFFFFFFll || method < 0 || method > 1)
    return MOCAT_ERR_ARG;
  hipLaunchKernelGGL(resample_kernel, dim3(1), dim3(NT), 0, stream, weights, n, key, method, from_log,
                     (uint64_t*)workspace, idx);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_resample_categorical(const float* log_weight, int64_t n, int64_t n_out, const uint32_t* key, int32_t* idx,
                               hipStream_t stream) {
  if (!log_weight || !key || !idx || n <= 0 || n_out <= 0) return MOCAT_ERR_ARG;
  if (n * n_out >= 0xFFFFFFFFll) return MOCAT_ERR_SIZE;
  const int64_t blocks = (n_out + 255) / 256;
  hipLaunchKernelGGL(categorical_kernel, dim3((unsigned)(blocks < 65535 ? blocks : 65535)), dim3(256), 0, stream,
                     log_weight, n, n_out, key, idx);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
