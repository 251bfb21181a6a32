// Tempered SMC sampler pieces (mocat/src/transport/smc.py):
//   gather_rows   resampled_ensemble_state = ensemble_state[resampled_indices] (smc.py:61-71)
//   smc_temper    one workgroup, FP64:
//     next temperature by mocat's regula falsi ("bisect", mocat/src/utils.py:205-237) on
//       f(b) = log_ess(lw - (b - b_t) lik) - log(ess_t * retain)   (smc.py:303-326)
//     or a fixed next temperature (temperature_schedule);
//     reweight lw_new = lw - (b' - b_t) lik (MetropolisedSMCSampler.log_weight, smc.py:367-373),
//     ESS and the log-normaliser increment lse(lw_new) - lse(lw) (TemperedSMCSampler.adapt,
//     smc.py:193-217), and the Robbins-Monro stepsize step with the weighted mean acceptance
//     (RMMetropolisedSMCSampler.adapt, smc.py:406-421).
#include "common.hpp"

using namespace mocat;

namespace {

constexpr int NT = 1024;
constexpr int NW = NT / kWave;

__global__ __launch_bounds__(256) void gather_kernel(const uint32_t* __restrict__ src, const int32_t* __restrict__ idx,
                                                     int64_t n_out, int64_t width, uint32_t* __restrict__ dst) {
  const int64_t total = n_out * width;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / width, k = e - j * width;
    dst[e] = src[(int64_t)idx[j] * width + k];
  }
}

struct Smem {
  double m[NW], s1[NW], s2[NW];
  double bc[8];
};

// (lse(w), lse(2w)) of w_i = lw_i - db * lik_i over the block, FP64, in two passes: the maximum m,
// then sum exp(w_i - m) and its square (one exponential per element); partial sums are combined in a
// fixed order (lanes by butterfly, waves in index order).
__device__ void lse_pair(const float* __restrict__ lw, const float* __restrict__ lik, int64_t n, double db, Smem& s,
                         double& lse, double& lse2) {
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  double mx = -INFINITY;
  int has_nan = 0;
  for (int64_t i = t; i < n; i += NT) {
    const double v = (double)lw[i] - db * (double)lik[i];
    if (isnan(v)) has_nan = 1;
    else mx = fmax(mx, v);
  }
  const bool nan_seen = __syncthreads_or(has_nan) != 0;
  mx = wave_max(mx);
  if (lane == 0) s.m[w] = mx;
  __syncthreads();
  double m = s.m[0];
  for (int i = 1; i < NW; ++i) m = fmax(m, s.m[i]);
  if (nan_seen || m == -INFINITY) {
    __syncthreads();  // every thread has read s.m before a later call rewrites it
    lse = nan_seen ? NAN : -INFINITY;
    lse2 = lse;
    return;
  }
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = t; i < n; i += NT) {
    const double e = exp((double)lw[i] - db * (double)lik[i] - m);
    s1 += e;
    s2 += e * e;
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  __syncthreads();
  if (lane == 0) {
    s.s1[w] = s1;
    s.s2[w] = s2;
  }
  __syncthreads();
  double a1 = 0.0, a2 = 0.0;
  for (int i = 0; i < NW; ++i) {
    a1 += s.s1[i];
    a2 += s.s2[i];
  }
  __syncthreads();  // every thread has read the partials before a later call rewrites them
  lse = m + log(a1);
  lse2 = 2.0 * m + log(a2);
}

struct TemperArgs {
  const float* lw;        // [n] current (resampled) log weights
  const float* lik;       // [n] likelihood potentials of the moved particles
  int64_t n;
  double beta;            // current temperature b_t
  double ess;             // ESS of lw (n after a resample)
  double retain;          // ess_threshold_retain
  double max_temperature;
  double tol;
  int max_iter;
  int adaptive;           // 1: regula falsi, 0: fixed next temperature `next_beta`
  double next_beta;
  const float* alpha;     // [k_alpha][n] acceptance probabilities of the moves, or NULL
  int k_alpha;
  double log_stepsize;    // current log stepsize
  double rm_stepsize;     // 0: no RM update
  double target;
  float* lw_out;          // [n]
  double* out;            // [8]: beta_new, ess_new, lse_new, lse_prev, log_stepsize_new, alpha_mean, iters, f_min
};

__global__ __launch_bounds__(NT) void temper_kernel(TemperArgs a) {
  __shared__ Smem s;
  const int t = threadIdx.x;
  double lse, lse2;
  const double logthr = log(a.ess * a.retain);
  auto f = [&](double b) {
    lse_pair(a.lw, a.lik, a.n, b - a.beta, s, lse, lse2);
    return (2.0 * lse - lse2) - logthr;
  };
  double bnew, fmin_abs = 0.0;
  int iters = 0;
  if (a.adaptive) {
    double b0 = a.beta, b1 = a.max_temperature;
    double e0 = f(b0), e1 = f(b1);
    const bool increasing = e1 > e0;
    while (true) {
      const bool done = fmin(fabs(e0), fabs(e1)) < a.tol || iters >= a.max_iter || (e0 < 0 && e1 < 0) ||
                        (e0 > 0 && e1 > 0) || isnan(e0) || isnan(e1);
      if (done) break;
      const double np = b0 - e0 * (b1 - b0) / (e1 - e0);
      const double ne = f(np);
      const bool replace_upper = increasing ? ne > 0 : ne < 0;
      if (replace_upper) {
        b1 = np;
        e1 = ne;
      } else {
        b0 = np;
        e0 = ne;
      }
      ++iters;
    }
    // bounds[argmin |evals|] (first on ties)
    bnew = fabs(e1) < fabs(e0) ? b1 : b0;
    fmin_abs = fmin(fabs(e0), fabs(e1));
  } else {
    bnew = a.next_beta;
  }
  const double db = bnew - a.beta;
  // previous lse (of lw) and the new weights
  double lse_prev, lse_prev2;
  lse_pair(a.lw, a.lik, a.n, 0.0, s, lse_prev, lse_prev2);
  double lse_new, lse_new2;
  lse_pair(a.lw, a.lik, a.n, db, s, lse_new, lse_new2);
  // weighted mean acceptance with weights exp(lw_new - max) (normalised)
  double amean = 0.0;
  if (a.alpha && a.k_alpha > 0) {
    double num = 0.0, den = 0.0;
    for (int64_t i = t; i < a.n; i += NT) {
      const double wi = exp((double)a.lw[i] - db * (double)a.lik[i] - lse_new);
      double am = 0.0;
      for (int k = 0; k < a.k_alpha; ++k) am += (double)a.alpha[(int64_t)k * a.n + i];
      am /= (double)a.k_alpha;
      num += wi * am;
      den += wi;
    }
    __shared__ double rnum[NW], rden[NW];
    num = wave_sum(num);
    den = wave_sum(den);
    if ((t & (kWave - 1)) == 0) {
      rnum[t / kWave] = num;
      rden[t / kWave] = den;
    }
    __syncthreads();
    double N = 0.0, D = 0.0;
    for (int i = 0; i < NW; ++i) {
      N += rnum[i];
      D += rden[i];
    }
    amean = N / D;
  }
  // written last so that lw_out may alias lw
  __syncthreads();
  for (int64_t i = t; i < a.n; i += NT) a.lw_out[i] = (float)((double)a.lw[i] - db * (double)a.lik[i]);
  if (t == 0) {
    a.out[0] = bnew;
    a.out[1] = exp(2.0 * lse_new - lse_new2);
    a.out[2] = lse_new;
    a.out[3] = lse_prev;
    a.out[4] = a.rm_stepsize != 0.0 ? a.log_stepsize + a.rm_stepsize * (amean - a.target) : a.log_stepsize;
    a.out[5] = amean;
    a.out[6] = (double)iters;
    a.out[7] = fmin_abs;
  }
}

}  // namespace

extern "C" {

int mocat_gather_rows(const void* src, const int32_t* idx, int64_t n_out, int64_t width, void* dst,
                      hipStream_t stream) {
  if (!src || !idx || !dst || n_out < 0 || width <= 0) return MOCAT_ERR_ARG;
  if (n_out == 0) return MOCAT_OK;
  const int64_t total = n_out * width;
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)(blocks < 65535 ? blocks : 65535)), dim3(256), 0, stream,
                     (const uint32_t*)src, idx, n_out, width, (uint32_t*)dst);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_smc_temper(const float* log_weight, const float* lik, int64_t n, double beta, double ess, double retain,
                     double max_temperature, double tol, int max_iter, int adaptive, double next_beta,
                     const float* alpha, int k_alpha, double log_stepsize, double rm_stepsize, double target,
                     float* log_weight_out, double* out, hipStream_t stream) {
  if (!log_weight || !lik || !log_weight_out || !out || n <= 0 || max_iter < 0 || k_alpha < 0) return MOCAT_ERR_ARG;
  TemperArgs a{log_weight, lik, n, beta, ess, retain, max_temperature, tol, max_iter, adaptive, next_beta,
               alpha, k_alpha, log_stepsize, rm_stepsize, target, log_weight_out, out};
  hipLaunchKernelGGL(temper_kernel, dim3(1), dim3(NT), 0, stream, a);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
