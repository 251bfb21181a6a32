/* mocat_amd C ABI: the drop-in boundary between mocat-shaped host code and the gfx950 kernels.
 *
 * Conventions (every entry point):
 *   - pointers are device pointers unless named host_*; sizes are explicit int64_t;
 *   - RNG keys are uint32_t[2] pairs (jax.random key layout), batched as [n_keys][2];
 *   - particle/chain arrays are row-major [count][dim] f32 unless stated;
 *   - the last argument is the hipStream_t the work is queued on (work is asynchronous);
 *   - the return value is a status: 0 ok, -1 invalid argument, -2 HIP launch/runtime error,
 *     -3 size limit exceeded. No entry point allocates device memory unless stated.
 *
 * The reference (mocat 0.2.6) is pure Python/JAX and has no C interface; each declaration cites
 * the reference function (file:line) whose computation it replaces.
 */
#ifndef MOCAT_AMD_H
#define MOCAT_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Returns the ABI revision of the library (not a status). */
int mocat_abi_version(void);

/* ---------------------------------------------------------------------------------------- RNG
 * jax.random.split / _random_bits / uniform / normal as called throughout mocat, e.g.
 * mocat/src/ssm/filtering.py:181 (split(key, n)), mocat/src/mcmc/standard_mcmc.py:120-122
 * (split + normal), mocat/src/mcmc/metropolis.py:60-61 (split + uniform).
 * Batched over n_keys keys: out is [n_keys][...]. */

/* split(keys[i], num) -> out[i][num][2]. */
int mocat_random_split(const uint32_t* keys, int64_t n_keys, int64_t num, uint32_t* out, hipStream_t stream);
/* _random_bits(keys[i], 32, (size,)) -> out[i][size]. size < 2^32 - 1. */
int mocat_random_bits(const uint32_t* keys, int64_t n_keys, int64_t size, uint32_t* out, hipStream_t stream);
/* uniform(keys[i], (size,), minval, maxval) -> out[i][size] (f32). */
int mocat_random_uniform(const uint32_t* keys, int64_t n_keys, int64_t size, float minval, float maxval,
                         float* out, hipStream_t stream);
/* normal(keys[i], (size,)) -> out[i][size] (f32). */
int mocat_random_normal(const uint32_t* keys, int64_t n_keys, int64_t size, float* out, hipStream_t stream);

/* --------------------------------------------------------------------------------- reductions
 * logsumexp over rows of w[batch][n] (f32 in, FP64 accumulation and results): lse[b] =
 * logsumexp(w[b]), lse2[b] = logsumexp(2 w[b]); either output may be NULL. log ESS = 2 lse - lse2
 * (mocat/src/metrics.py:69-78; jax.scipy.special.logsumexp at mocat/src/transport/smc.py:160,212-215). */
int mocat_logsumexp(const float* w, int64_t batch, int64_t n, double* lse, double* lse2, hipStream_t stream);

/* ------------------------------------------------------------------------ logistic regression
 * Fused potential + adjoint of the Bayesian logistic-regression Scenario (user Scenario pattern,
 * README.md:35-49; gradients replace jax.grad in mocat/src/core.py:204-230):
 *   prior = 0.5*prec*|t|^2 + 0.5*d*(log 2pi - log prec), lik = sum softplus(x.t) - y*x.t,
 *   grad_lik = X^T (sigmoid(X t) - y). X is [n][d] f32 row-major, y [n] f32, d <= 128.
 * precision: 0 = FP32 vector FMAs (reads X, y), 1 = bf16x3 MFMA (reads the packed buffer). */

/* Bytes needed for the packed (bf16 hi/lo, MFMA operand order) copy of X and y. */
int mocat_lr_packed_bytes(int64_t n, int64_t d, int64_t* bytes);
/* Writes the packed copy of X, y into `packed` (device, mocat_lr_packed_bytes bytes). */
int mocat_lr_pack(const float* X, const float* y, int64_t n, int64_t d, void* packed, hipStream_t stream);
/* For C chains theta[C][d]: prior_pot[C], lik_pot[C], grad_prior[C][d], grad_lik[C][d]
 * (any output may be NULL). Replaces vmap(Scenario.*_potential_and_grad), mocat/src/core.py:204-230. */
int mocat_lr_potential_grad(const float* X, const float* y, const void* packed, int64_t n, int64_t d,
                            int precision, float prior_prec, float temperature, const float* theta, int64_t C,
                            float* prior_pot, float* lik_pot, float* grad_prior, float* grad_lik,
                            hipStream_t stream);

/* n_iter iterations of MCMCSampler.update for C independent chains of Underdamped (HMC when
 * friction = inf; MALA when also leapfrog_steps = 1) on the logistic-regression scenario:
 * always (mocat/src/mcmc/standard_mcmc.py:107-123), leapfrog (mocat/src/utils.py:108-146),
 * proposal/acceptance (standard_mcmc.py:125-153), correction (mocat/src/mcmc/metropolis.py:48-161).
 * correction: 0 uncorrected, 1 Metropolis, 2 RMMetropolis (log Robbins-Monro, step iter^-exp*scale).
 * State arrays (in/out): theta, momenta, grad [C][d]; potential, prior_pot, lik_pot, stepsize,
 * alpha [C]; keys [C][2]. iter0 = extra.iter before the first iteration. Optional traces
 * (NULL to skip): trace_value/trace_momenta [n_iter][C][d], trace_potential/alpha/stepsize
 * [n_iter][C]. */
int mocat_hmc_lr_iterate(const float* X, const float* y, const void* packed, int64_t n, int64_t d, int precision,
                         float prior_prec, float temperature, int64_t C, float* theta, float* momenta, float* grad,
                         float* potential, float* prior_pot, float* lik_pot, float* stepsize, float* alpha,
                         uint32_t* keys, int64_t iter0, int n_iter, int leapfrog_steps, float friction,
                         int correction, float target, float rm_scale, float rm_neg_exponent, double adapt_cut_off,
                         float* trace_value, float* trace_potential, float* trace_alpha, float* trace_stepsize,
                         float* trace_momenta, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MOCAT_AMD_H */
