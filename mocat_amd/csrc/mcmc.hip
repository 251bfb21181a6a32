// The Underdamped (HMC / MALA / underdamped Langevin) iteration for scenarios without a fused kernel:
// potentials and gradients come from the caller between launches (for example a user Scenario
// written in torch), every other step of the iteration runs here with the same key tree and the same
// arithmetic as the fused iteration kernel in lr.hip.
//
//   hmc_begin    always (mocat/src/mcmc/standard_mcmc.py:107-123): p <- -p, (key, sub) = split(key),
//                OU refresh of p from normal(sub, (d,)); then the proposal's keys = split(key, L + 1),
//                key <- keys[0] and keys[1..L] handed out for the L leapfrog potentials
//                (standard_mcmc.py:125-137); the refreshed momenta and their kinetic energy are kept
//                for a rejection and for the acceptance probability.
//   hmc_kick     one half step of the leapfrog (mocat/src/utils.py:117-146): p <- p - h/2 grad, and
//                with drift also x <- x + h p.
//   hmc_correct  Metropolis / RMMetropolis (mocat/src/mcmc/metropolis.py:48-70, 118-161) with
//                Underdamped's acceptance probability (standard_mcmc.py:139-153): on acceptance the
//                proposal becomes the state with momenta -p', on rejection the momenta are the
//                refreshed ones; then the log Robbins-Monro stepsize update.
// One wave per chain for begin / correct (the kinetic energy is a wave reduction), a flat grid for
// the kick.
#include "common.hpp"

using namespace mocat;

namespace {

constexpr int kThreads = 256;
constexpr int kChainsPerBlock = kThreads / kWave;
constexpr float kLog2Pi = 1.83787706640934548356f;

__global__ __launch_bounds__(kThreads) void hmc_begin_kernel(int64_t C, int d, int leapfrog_steps, float friction,
                                                             const float* __restrict__ step, float* __restrict__ mom,
                                                             float* __restrict__ mom0, float* __restrict__ kin0,
                                                             uint32_t* __restrict__ keys,
                                                             uint32_t* __restrict__ step_keys) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * kChainsPerBlock + threadIdx.x / kWave;
  if (c >= C) return;
  const Key k{keys[2 * c], keys[2 * c + 1]};
  const Key nk = split_at(k, 2, 0), sk = split_at(k, 2, 1);
  const uint32_t L1 = (uint32_t)leapfrog_steps + 1u;
  if (step_keys)
    for (int j = lane; j < leapfrog_steps; j += kWave) {
      const Key q = split_at(nk, L1, (uint32_t)j + 1u);
      step_keys[2 * (c * leapfrog_steps + j)] = q.a;
      step_keys[2 * (c * leapfrog_steps + j) + 1] = q.b;
    }
  const float hh = step[c];
  const float decay = __expf(-friction * hh);
  const float noise = sqrtf(1.f - __expf(-2.f * friction * hh));
  const uint32_t hd = ((uint32_t)d + 1u) >> 1;  // threefry blocks of normal(sub, (d,))
  float* p = mom + c * d;
  float* p0 = mom0 + c * d;
  float s2 = 0.f;
  for (uint32_t i = lane; i < hd; i += kWave) {
    uint32_t w0, w1;
    bits_pair(sk, (uint32_t)d, i, w0, w1);
    const float a = -p[i] * decay + noise * normal_from_bits(w0);
    p[i] = a;
    p0[i] = a;
    s2 += a * a;
    if (i + hd < (uint32_t)d) {
      const float b = -p[i + hd] * decay + noise * normal_from_bits(w1);
      p[i + hd] = b;
      p0[i + hd] = b;
      s2 += b * b;
    }
  }
  s2 = wave_sum(s2);
  if (lane == 0) {
    kin0[c] = 0.5f * s2 + 0.5f * (float)d * kLog2Pi;
    const Key pk = split_at(nk, L1, 0);
    keys[2 * c] = pk.a;
    keys[2 * c + 1] = pk.b;
  }
}

__global__ __launch_bounds__(kThreads) void hmc_kick_kernel(int64_t C, int d, float* __restrict__ x,
                                                            float* __restrict__ mom, const float* __restrict__ grad,
                                                            const float* __restrict__ step, int drift) {
  const int64_t total = C * (int64_t)d;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kThreads) {
    const float hh = step[e / d];
    const float ph = mom[e] - (hh / 2.f) * grad[e];
    mom[e] = ph;
    if (drift) x[e] = x[e] + hh * ph;
  }
}

struct CorrectArgs {
  int64_t C;
  int d;
  int64_t iter;  // the iteration number after this one (1 for the first), as the RM schedule counts
  int correction;
  float target, rm_scale, rm_neg_exponent;
  double adapt_cut_off;
  float *x, *grad, *pot, *prior, *lik;                    // state, replaced by the proposal on acceptance
  const float *x1, *grad1, *pot1, *prior1, *lik1;         // proposal
  float* mom;                                             // proposal momenta p' in, the state's momenta out
  const float* mom0;                                      // refreshed momenta (hmc_begin)
  const float* kin0;
  float *step, *alpha;
  uint32_t* keys;
};

__global__ __launch_bounds__(kThreads) void hmc_correct_kernel(CorrectArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * kChainsPerBlock + threadIdx.x / kWave;
  if (c >= a.C) return;
  const int d = a.d;
  float* p = a.mom + c * d;
  float s2 = 0.f;
  for (int k = lane; k < d; k += kWave) s2 += p[k] * p[k];
  s2 = wave_sum(s2);
  const float kin1 = 0.5f * s2 + 0.5f * (float)d * kLog2Pi;  // |-p'|^2 == |p'|^2
  const float pot0 = a.pot[c], pot1 = a.pot1[c];
  Key key{a.keys[2 * c], a.keys[2 * c + 1]};
  float al = 1.f;
  bool acc = true;
  if (a.correction > 0) {
    al = fminf(1.f, __expf(-pot1 + pot0 - kin1 + a.kin0[c]));
    al = isnan(al) ? 0.f : al;
    const Key nk = split_at(key, 2, 0), sk = split_at(key, 2, 1);
    key = nk;
    acc = uniform_from_bits(bits_at(sk, 1, 0), 0.f, 1.f) < al;
  }
  const int64_t o = c * d;
  for (int k = lane; k < d; k += kWave) {
    if (acc) {
      p[k] = -p[k];
      a.x[o + k] = a.x1[o + k];
      a.grad[o + k] = a.grad1[o + k];
    } else {
      p[k] = a.mom0[o + k];
    }
  }
  if (lane == 0) {
    a.alpha[c] = al;
    a.keys[2 * c] = key.a;
    a.keys[2 * c + 1] = key.b;
    if (acc) {
      a.pot[c] = pot1;
      a.prior[c] = a.prior1[c];
      a.lik[c] = a.lik1[c];
    }
    if (a.correction == 2 && (double)a.iter <= a.adapt_cut_off) {
      const float rm = a.rm_scale * powf((float)a.iter, -a.rm_neg_exponent);
      a.step[c] = expf(logf(a.step[c]) + rm * (al - a.target));
    }
  }
}

unsigned chain_blocks(int64_t C) { return (unsigned)((C + kChainsPerBlock - 1) / kChainsPerBlock); }

}  // namespace

extern "C" {

int mocat_hmc_begin(int64_t C, int d, int leapfrog_steps, float friction, const float* step, float* mom, float* mom0,
                    float* kin0, uint32_t* keys, uint32_t* step_keys, hipStream_t stream) {
  if (C <= 0 || d <= 0 || leapfrog_steps < 0 || !step || !mom || !mom0 || !kin0 || !keys) return MOCAT_ERR_ARG;
  if (C > ((int64_t)1 << 31) || (int64_t)C * d >= ((int64_t)1 << 40)) return MOCAT_ERR_SIZE;
  hipLaunchKernelGGL(hmc_begin_kernel, dim3(chain_blocks(C)), dim3(kThreads), 0, stream, C, d, leapfrog_steps,
                     friction, step, mom, mom0, kin0, keys, step_keys);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_hmc_kick(int64_t C, int d, float* x, float* mom, const float* grad, const float* step, int drift,
                   hipStream_t stream) {
  if (C <= 0 || d <= 0 || !mom || !grad || !step || (drift && !x)) return MOCAT_ERR_ARG;
  const int64_t total = C * (int64_t)d;
  const int64_t blocks = (total + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(hmc_kick_kernel, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(kThreads), 0, stream, C,
                     d, x, mom, grad, step, drift);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

int mocat_hmc_correct(int64_t C, int d, int64_t iter, int correction, float target, float rm_scale,
                      float rm_neg_exponent, double adapt_cut_off, float* x, float* grad, float* pot, float* prior,
                      float* lik, const float* x1, const float* grad1, const float* pot1, const float* prior1,
                      const float* lik1, float* mom, const float* mom0, const float* kin0, float* step, float* alpha,
                      uint32_t* keys, hipStream_t stream) {
  if (C <= 0 || d <= 0 || correction < 0 || correction > 2 || !x || !grad || !pot || !prior || !lik || !x1 ||
      !grad1 || !pot1 || !prior1 || !lik1 || !mom || !mom0 || !kin0 || !step || !alpha || !keys)
    return MOCAT_ERR_ARG;
  CorrectArgs a{C, d, iter, correction, target, rm_scale, rm_neg_exponent, adapt_cut_off, x, grad, pot, prior, lik,
                x1, grad1, pot1, prior1, lik1, mom, mom0, kin0, step, alpha, keys};
  hipLaunchKernelGGL(hmc_correct_kernel, dim3(chain_blocks(C)), dim3(kThreads), 0, stream, a);
  MOCAT_LAUNCH_CHECK();
  return MOCAT_OK;
}

}  // extern "C"
