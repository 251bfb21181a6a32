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
int mocat_abi_version(void); /* this header: 5 */

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
/* randint(keys[i], (size,), minval, maxval) -> out[i][size] (int32), the draw behind
 * random.choice(key, n, shape) with replacement in mocat/src/metrics.py:103 (KSD minibatches). */
int mocat_random_randint(const uint32_t* keys, int64_t n_keys, int64_t size, int32_t minval, int32_t maxval,
                         int32_t* out, hipStream_t stream);

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
 * precision: 0 = FP32 vector FMAs (reads X, y), 1 = bf16x3 MFMA (reads the packed buffer),
 * 2 = diagonal Gaussian likelihood (X = mean[d], y = precisions[d]), 3 = FP64 vector FMAs with f32
 * inputs and outputs (reads X, y; the reference for the error of 0 and 1). */

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

/* The same iteration in three pieces for scenarios without a fused kernel (the caller evaluates the
 * potentials and gradients between them, e.g. a user Scenario's jax.grad-derived
 * *_potential_and_grad, mocat/src/core.py:204-230), with the fused kernel's key tree and arithmetic.
 * Per iteration: mocat_hmc_begin, then per leapfrog step kick(drift = 1) on the proposal, the
 * caller's evaluation at the new position, kick(drift = 0); then mocat_hmc_correct.
 * begin: Underdamped.always (mocat/src/mcmc/standard_mcmc.py:107-123) and the proposal's key split
 * (:125-137): momenta [C][d] refreshed in place and copied to mom0, kin0 [C] their kinetic energy,
 * keys [C][2] advanced to split(key', L + 1)[0], step_keys [C][L][2] (or NULL) = the leapfrog
 * potentials' keys split(key', L + 1)[1..L]. */
int mocat_hmc_begin(int64_t C, int d, int leapfrog_steps, float friction, const float* stepsize, float* momenta,
                    float* mom0, float* kin0, uint32_t* keys, uint32_t* step_keys, hipStream_t stream);
/* Half a leapfrog step (mocat/src/utils.py:117-146): momenta -= stepsize/2 * grad; with drift,
 * x += stepsize * momenta. */
int mocat_hmc_kick(int64_t C, int d, float* x, float* momenta, const float* grad, const float* stepsize, int drift,
                   hipStream_t stream);
/* Metropolis / RMMetropolis correction of the proposal (x1, grad1, pot1, prior1, lik1; momenta = p')
 * against the state (x, grad, pot, prior, lik), mocat/src/mcmc/metropolis.py:48-161 with
 * Underdamped.acceptance_probability (standard_mcmc.py:139-153). iter = extra.iter after this
 * iteration; correction and the RM arguments as mocat_hmc_lr_iterate. */
int mocat_hmc_correct(int64_t C, int d, int64_t iter, int correction, float target, float rm_scale,
                      float rm_neg_exponent, double adapt_cut_off, float* x, float* grad, float* pot, float* prior,
                      float* lik, const float* x1, const float* grad1, const float* pot1, const float* prior1,
                      const float* lik1, float* momenta, const float* mom0, const float* kin0, float* stepsize,
                      float* alpha, uint32_t* keys, hipStream_t stream);

/* ------------------------------------------------------------------------- particle filters
 * run_particle_filter_for_marginals with a BootstrapFilter (mocat/src/ssm/filtering.py:255-324,
 * 142-170) for many independent filters of uneven n and T in one launch (one workgroup per filter).
 * model 0: TimeHomogenousLinearGaussian (linear_gaussian.py:46-51,86-94,118-128), params
 *   [m0(d), L0(dxd), A(dxd), LQ(dxd), H(dyxd), SR(dyxdy), nlz] with SR = inv(chol(R)) and
 *   nlz = (dy log 2pi - log det R^-1)/2 as mocat's reset_covariance (mocat/src/utils.py:247-262);
 * model 1: StochasticVolatility (stochastic_volatility.py:15-108), d = dy in 1..4; d = 1 params
 *   [m0, sqrt(P0), phi, sigma, 1/beta], d > 1 params [m0(d), L0(dxd), A(dxd), LQ(dxd), R^-1/2(d)]
 *   with L0, LQ the Cholesky factors of P0 and Q and R the likelihood covariance's diagonal;
 *   observation times enter through step_scale (below);
 * model 2: TimeHomogenousLinearGaussian under OptimalNonLinearGaussianParticleFilter
 *   (mocat/src/ssm/nonlinear_gaussian.py:134-281), params [m0(d), H(dyxd), K0(dxdy), C0(dxd),
 *   A(dxd), K(dxdy), S(dxd), WC(dyxdy), nlzw] as its startup (:154-184); initial weights zero;
 *   every particle's noise from normal(split(key, n)[0], (n, d)); n*d < 2^32.
 * Key tree per filter: (key, sub) = split(key); particles from split(sub, n); step t uses
 * split(key, T-1)[t-1] for resampling and split(., n) for the transitions (filtering.py:264-303).
 * resampling: 0 systematic, 1 multinomial, 2 categorical (mocat's Gumbel-max, n*n < 2^32 - 1);
 * resample when ESS_{t-1} < ess_threshold * n. Outputs (optional unless noted): trajectories at
 * out_offset ([T][n][d], [T][n], ancestors [T][n]), ess / cumulative log-likelihood at t_offset [T],
 * out_loglik[n_filters] (required, FP64). Workspaces are sized by the sum of n over filters.
 * out_grad (model 1 with d = 1 only, else NULL): [n_filters][3] d log Z / d(phi, sigma, beta) of the
 * fixed-resampling-index estimator by forward tangents (workspaces tan_ws0/1: [sum n][5] floats). */
typedef struct {
  int64_t n;             /* particles */
  int64_t T;             /* observations y_0..y_{T-1} */
  int64_t y_offset;      /* floats into y */
  int64_t state_offset;  /* particle slots into the workspaces */
  int64_t out_offset;    /* particle-steps into the trajectory outputs, or -1 */
  int64_t t_offset;      /* steps into out_ess / out_loglik_t */
  uint32_t key[2];       /* random_key of the filter */
  int32_t param_offset;  /* floats into params */
  int32_t flags;         /* bit 0 (MOCAT_PF_LARGE): run by mocat_pf_run_large, skipped by mocat_pf_run;
                            bits 1-3: MOCAT_PF_RESUME, MOCAT_PF_STEP_KEY, MOCAT_PF_INIT_KEY below */
} mocat_pf_filter;

#define MOCAT_PF_LARGE 1
/* run_particle_filter_for_marginals with initial_sample (mocat/src/ssm/filtering.py:264-278): the
 * filter's particles and log-weights at t = 0 are read from x_ws0 / lw_ws0 at state_offset (the caller
 * writes them) instead of drawn, y_0 is not read, step t uses split(key, T-1)[t-1], and the
 * log-likelihood starts at 0 (it estimates log p(y_1:T-1 | the given sample)). Not with out_grad. */
#define MOCAT_PF_RESUME 2
/* mocat's one-step functions (filtering.py:173-253): MOCAT_PF_STEP_KEY makes the filter key itself the
 * key of step 1 (propagate_particle_filter, with T = 2 and MOCAT_PF_RESUME); MOCAT_PF_INIT_KEY draws
 * the initial particles from split(key, n) without the (key, sub) split (initiate_particles, T = 1). */
#define MOCAT_PF_STEP_KEY 4
#define MOCAT_PF_INIT_KEY 8

int mocat_pf_param_count(int model, int d, int dy, int64_t* count);
/* *compiled = 1 when mocat_pf_run / mocat_pf_run_large have kernels for (model, d, dy) (model 1:
 * d = dy <= 4; models 0 and 2: the (d, dy) pairs (1,1) (2,1) (2,2) (3,3) (4,1) (4,2) (4,4) (5,5) (8,4)
 * (8,8)), else 0. No counterpart in the reference (mocat traces any shape); its hosts use it to choose
 * between the fused filters and the hook loop. */
int mocat_pf_compiled_shape(int model, int d, int dy, int* compiled);
/* host_filters and filters hold the same n_filters descriptors: host_filters in host memory (read by
 * the argument checks, not retained after return), filters in device memory (read by the kernel).
 * Every filter not flagged MOCAT_PF_LARGE needs n >= 1, T >= 1, non-negative offsets and, when
 * out_ess or out_loglik_t is given, t_offset >= 0.
 * step_scale (NULL: unit time steps) holds, per row of y (index y_offset / dy + t), the noise scale of
 * the step into row t for the stochastic-volatility model (model 1): sqrt(t_t - t_{t-1}), mocat's
 * transition_sample (mocat/src/ssm/scenarios/stochastic_volatility.py:86-93); the linear-Gaussian
 * models are time homogeneous and do not read it. */
int mocat_pf_run(int model, int d, int dy, const mocat_pf_filter* host_filters, const mocat_pf_filter* filters,
                 int n_filters, const float* params,
                 const float* y, const float* step_scale, float* x_ws0, float* x_ws1, float* lw_ws0, float* lw_ws1, uint64_t* csum_ws,
                 int32_t* anc_ws, int resampling, float ess_threshold, float* out_value, float* out_lw,
                 int32_t* out_anc, float* out_ess, double* out_loglik_t, double* out_loglik, float* tan_ws0,
                 float* tan_ws1, double* out_grad, hipStream_t stream);

/* The same filters (same run_particle_filter_for_marginals semantics, outputs and workspaces as
 * mocat_pf_run) for those flagged MOCAT_PF_LARGE, spread over the whole GPU: per time step three
 * grid-wide launches over 1024-particle chunks, all flagged filters batched in each (the previous
 * step's statistics, resampling decision and quantised weight totals; exact prefix sums; ancestor
 * search, propagation and weighting). For the same log-weights and key the ancestors equal mocat_pf_run's. host_filters is a
 * host array (the launch sequence depends on n and T); workspace holds
 * mocat_pf_large_workspace_bytes bytes, including T step-key slots per flagged filter (independent of
 * t_offset, which only places out_ess / out_loglik_t and must be >= 0 when either is given).
 * launch_mode 0 queues the per-step launches on the stream; 1 captures them into one HIP graph,
 * launches it on the stream and returns after it has run (when the stream is already being captured
 * by the caller, the launches join the caller's graph instead); 2 runs all steps in one persistent
 * kernel (at most four workgroups per CU, all resident) with a grid barrier between the phases; a
 * barrier that waits more than 2^22 polls (or the count in the environment variable
 * MOCAT_PF_BARRIER_POLLS) gives up, and every flagged filter's out_loglik (and out_grad) is then NaN. Results do not depend on the mode. The workspace also holds the step's split
 * keys (8 bytes per particle, rounded up to whole chunks).
 * Calls on one stream may follow mocat_pf_run with the same arguments: the two paths touch disjoint
 * filters. */
int mocat_pf_large_workspace_bytes(const mocat_pf_filter* host_filters, int n_filters, int64_t* bytes);
int mocat_pf_run_large(int model, int d, int dy, const mocat_pf_filter* host_filters, int n_filters,
                       const float* params,
                       const float* y, const float* step_scale, float* x_ws0, float* x_ws1, float* lw_ws0, float* lw_ws1, uint64_t* csum_ws,
                       int resampling, float ess_threshold, float* out_value, float* out_lw, int32_t* out_anc,
                       float* out_ess, double* out_loglik_t, double* out_loglik, float* tan_ws0, float* tan_ws1,
                       double* out_grad, void* workspace, int64_t workspace_bytes, int launch_mode,
                       hipStream_t stream);

/* The batched per-step hooks of mocat's particle filters for the compiled models (model, d, dy and
 * params as mocat_pf_run), one call per step over n particles, with the fused filters' noise draws:
 * phase 0 = ParticleFilter.initial_sample_and_weight_vectorised (mocat/src/ssm/filtering.py:65-78):
 *   x[i] = initial sample from keys[i], log_weight[i] = -likelihood(x[i], y) (y may be NULL: zeros);
 *   model 2 (optimal proposal) conditions on y (required) from row i of normal(keys[0], (n, d)) with
 *   log_weight 0 (mocat/src/ssm/nonlinear_gaussian.py:198-224);
 * phase 1 = propose_and_intermediate_weight_vectorised (filtering.py:124-139, BootstrapFilter
 *   :142-170; optimal proposal nonlinear_gaussian.py:264-281): x[i] from x_prev[i] and keys[i] (model
 *   2: keys[0]), log_weight[i] = the intermediate log-weight at y;
 * phase 2 = StateSpaceModel.likelihood_potential (mocat/src/ssm/ssm.py:73-81) of x_prev[i] at y into
 *   log_weight[i] (models 0 and 1; x and keys unused).
 * x_prev, x [n][d], y [dy], keys [n][2] (model 2: keys[0] only), log_weight [n]; step_scale: the
 * step's noise scale for model 1 in phase 1 (sqrt(t_new - t_previous); 1 for unit steps). Models 0
 * and 2 at a (d, dy) that mocat_pf_compiled_shape does not list run a runtime-dimension kernel (d,
 * dy <= 16, the same formulas and noise draws): the host runs such a filter's time loop over these hooks,
 * mocat_gather_rows, mocat_resample / mocat_resample_categorical and mocat_logsumexp. */
int mocat_pf_hook(int model, int d, int dy, int phase, const float* params, const float* x_prev, const float* y,
                  const uint32_t* keys, float step_scale, int64_t n, float* x, float* log_weight,
                  hipStream_t stream);

/* Ancestor indices idx[n] from non-negative weights (from_log = 0) or log-weights (from_log = 1,
 * weights exp(w - max w)): method 0 systematic (u = uniform(key)), 1 multinomial
 * (u = uniform(key, (n,))); exact fixed-point cumulative weights. key is a device uint32[2];
 * workspace holds n uint64. New resamplers beside mocat's _resample (filtering.py:196-199). */
int mocat_resample(const float* weights, int64_t n, int method, int from_log, const uint32_t* key, void* workspace,
                   int32_t* idx, hipStream_t stream);
/* mocat_resample without a workspace argument and without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). The prefix sums go to a library-owned device array of
 * 2^20 uint64 (8 MiB per GPU, held until the process ends): n > 2^20 returns -3 (size limit). All
 * calls on one GPU share that array, so calls that overlap on one GPU (different streams, or other
 * threads) race; calls ordered on one stream do not. Same results as mocat_resample. */
int mocat_resample_noalloc(const float* weights, int64_t n, int method, int from_log, const uint32_t* key,
                           int32_t* idx, hipStream_t stream);
/* random.categorical(key, log_weight, shape=(n_out,)) by Gumbel-max: mocat's _resample
 * (mocat/src/ssm/filtering.py:196-199) and SMCSampler.resample (mocat/src/transport/smc.py:61-71). */
int mocat_resample_categorical(const float* log_weight, int64_t n, int64_t n_out, const uint32_t* key, int32_t* idx,
                               hipStream_t stream);

/* ---- Backward simulation, full resampling (replaces mocat/src/ssm/backward.py:20-39 and the scan
 * body of backward_simulation_full, :226-237). One backward step for m trajectories:
 * x_out[k] = particles[idx[k]], idx[k] = argmax_j(gumbel(split(key, m)[k], (n,))[j] + log_weight[j]
 * - U(particles[j] -> x_next[k])) (first maximum), U the transition potential of model 0 (linear
 * Gaussian: compiled for d in {1,2,3,4,5,8}, a runtime-dimension kernel for any other d <= 16; params
 * A[d*d], inv(cholesky(Q))[d*d], normaliser) or model 1
 * (stochastic volatility, d = 1, unit steps; params phi, 1/sigma, normaliser). idx may be NULL.
 * particles [n][d], x_next and x_out [m][d]. */
int mocat_ffbsi_step(int model, int d, const float* params, const float* particles, const float* log_weight,
                     int64_t n, const float* x_next, int64_t m, const uint32_t* key, int32_t* idx, float* x_out,
                     hipStream_t stream);
/* ---- Backward simulation by rejection (replaces mocat/src/ssm/backward.py:42-194, one step of the
 * scan in backward_simulation_rejection, :197-207): initial categorical proposals, accept/reject
 * rounds against a shared density bound (raised by the inflation factor when exceeded, which sends
 * every trajectory back to proposing), at most max_rejections - 1 rounds, then full resampling for
 * the trajectories still waiting. evals (one int64 on the device) receives the step's number of
 * transition-density evaluations. Models and params as mocat_ffbsi_step (model 0 at any d <= 16);
 * workspace: mocat_ffbsi_rejection_workspace_bytes(m). */
int mocat_ffbsi_rejection_workspace_bytes(int64_t m, int64_t* bytes);
int mocat_ffbsi_rejection_step(int model, int d, const float* params, const float* particles, const float* log_weight,
                               int64_t n, const float* x_next, int64_t m, const uint32_t* key, int max_rejections,
                               float init_bound, float bound_inflation, void* workspace, int64_t workspace_bytes,
                               int32_t* idx, float* x_out, long long* evals, hipStream_t stream);

/* ----------------------------------------------------------------------- ensemble Kalman filter
 * run_particle_filter_for_marginals with EnsembleKalmanFilter (mocat/src/ssm/nonlinear_gaussian.py:
 * 284-350, filtering.py:255-324) for TimeHomogenousLinearGaussian (transition_function = A x):
 * params [m0(d), H(dyxd), K0(dxdy), C0(dxd), A(dxd), LQ(dxd), LR(dyxdy)] (K0, C0 as the filter's
 * startup :297-310, LQ/LR = cholesky of Q/R); y [T][dy]; key: device u32[2]. With
 * (rest, sub) = split(key): x_0 from normal(split(sub, n)[0], (n, d)); step t from
 * (k1, k2) = split(split(split(rest, T-1)[t-1], n)[0]), normal(k1, (n, d)) the forecast noise and
 * normal(k2, (n, dy)) the observation perturbation (filtering.py:267-294); gain from the forecast
 * ensemble's covariance (ddof 1) in FP64. x_out [n][d] final ensemble (required); out_value [T][n][d] and
 * out_gain [T-1][d][dy] optional. n >= 2, n * max(d, dy) < 2^32. Two launches per step. */
int mocat_enkf_param_count(int d, int dy, int64_t* count);
int mocat_enkf_workspace_bytes(int64_t n, int d, int64_t* bytes);
int mocat_enkf_run(int d, int dy, const float* params, const float* y, int64_t T, int64_t n, const uint32_t* key,
                   float* x_out, void* workspace, int64_t workspace_bytes, float* out_value, float* out_gain,
                   hipStream_t stream);
/* The EnKF's member updates at any (d, dy) up to 16 (mocat_enkf_run runs the (d, dy) pairs that
 * mocat_pf_compiled_shape lists for model 0), one launch per call, params as mocat_enkf_run and z the
 * rows of normal(key, (n, d)) (phase 2: (n, dy)); key a device uint32[2]:
 *   phase 0 (initial_sample, nonlinear_gaussian.py:312-324): x_out = m0 + K0 (y - H m0) + z @ C0;
 *   phase 1 (the forecast of :326-350): x_out = A x_in + LQ z;
 *   phase 2 (the analysis): x_out = x_in + K (y - (H x_in + LR z)) with the gain K [d][dy].
 * A host runs the filter as phase 0, then per step phase 1, the gain from the forecast ensemble's
 * covariance (FP64), and phase 2 (mocat_amd/ssm/filtering.py, _run_enkf_hooks). n * max(d, dy) < 2^32. */
int mocat_enkf_hook(int phase, int d, int dy, const float* params, const float* x_in, const uint32_t* key, int64_t n,
                    const float* y, const float* gain, float* x_out, hipStream_t stream);

/* -------------------------------------------------------------------------------- tempered SMC
 * dst[j][:] = src[idx[j]][:] for 4-byte elements, width per row: ensemble_state[resampled_indices]
 * (mocat/src/transport/smc.py:68). */
int mocat_gather_rows(const void* src, const int32_t* idx, int64_t n_out, int64_t width, void* dst,
                      hipStream_t stream);
/* One adapt step of (RM)MetropolisedSMCSampler in FP64 (mocat/src/transport/smc.py:193-217,
 * 303-326, 406-421; regula falsi of mocat/src/utils.py:205-237): next temperature (adaptive = 1,
 * ESS target ess * retain) or next_beta; log_weight_out = log_weight - (b' - beta) lik; out[8] =
 * {b', ESS', lse(new), lse(old), new log stepsize, weighted mean alpha, iterations, min |f|}.
 * alpha: [k_alpha][n] acceptance probabilities of the moves (NULL: no RM update). */
int mocat_smc_temper(const float* log_weight, const float* lik, int64_t n, double beta, double ess, double retain,
                     double max_temperature, double tol, int max_iter, int adaptive, double next_beta,
                     const float* alpha, int k_alpha, double log_stepsize, double rm_stepsize, double target,
                     float* log_weight_out, double* out, hipStream_t stream);

/* ------------------------------------------------------------------------------------- SVGD
 * One SVGD.update with the Gaussian kernel and adagrad(step, momentum)
 * (mocat/src/transport/svgd.py:18-32,122-146; kernels.py:89-102): x (in/out), gsq and mom
 * (adagrad state, in/out) [n][d]; x_in and grad are the positions and potential gradients the Stein
 * direction is computed from; phi_out optional. d <= 128. workspace: mocat_svgd_workspace_bytes.
 * The kernel bandwidth is read from bandwidth_dev (a device float, e.g. the output of the median
 * update, so the iteration needs no host round trip) when it is not NULL, else from bandwidth. */
int mocat_svgd_workspace_bytes(int64_t n, int64_t d, int js_count, int64_t* bytes);
int mocat_svgd_step(const float* x_in, const float* grad, int64_t n, int64_t d, float bandwidth,
                    const float* bandwidth_dev, int js_count, void* workspace, float step, float momentum, float* x,
                    float* gsq, float* mom, float* phi_out, hipStream_t stream);
/* Diagonal Gaussian mixture: potential[n] = -logsumexp_k(logc_k - 0.5 sum_j lam_kj (x_j - mu_kj)^2),
 * grad[n][d] (GaussianMixture, mocat/src/scenarios/toy_examples.py:54-116). K <= 8. */
int mocat_gmm_potential_grad(const float* x, int64_t n, int64_t d, int K, const float* means, const float* lam,
                             const float* logc, float* potential, float* grad, hipStream_t stream);
/* Gaussian mixture with full covariances: potential[n] = -logsumexp_k(logc_k - 0.5 r_k^T S_k r_k),
 * r_k = x - mu_k, grad[n][d] = sum_k softmax_k S_k r_k; S [K][d][d] symmetric (the precision P_k, or
 * L_k^T L_k with L_k = cholesky(P_k) for mocat's component potential as written,
 * mocat/src/scenarios/toy_examples.py:95-100). K <= 8, d <= 128. */
int mocat_gmm_full_potential_grad(const float* x, int64_t n, int64_t d, int K, const float* means,
                                  const float* S, const float* logc, float* potential, float* grad,
                                  hipStream_t stream);
/* d2[n][n] = |x_i - x_j|^2 (diagonal 0), the distance matrix of l2_distance_matrix
 * (mocat/src/utils.py:437-439) squared. workspace: mocat_svgd_workspace_bytes(n, d, 1). */
int mocat_pairwise_sqdist(const float* x, int64_t n, int64_t d, void* workspace, float* d2, hipStream_t stream);
/* Exact rank0-th and rank1-th smallest of m non-negative floats (radix select): the median of
 * median_bandwidth_update (mocat/src/kernels.py:220-224). Results (bit patterns) in workspace[0..1];
 * workspace: 516 uint32. */
int mocat_select2(const float* v, int64_t m, int64_t rank0, int64_t rank1, uint32_t* workspace, hipStream_t stream);
/* partial[b] = sum of sqrt(d2) over a grid-stride share of the m entries (FP64), for
 * mean_bandwidth_update (mocat/src/kernels.py:227-229). */
int mocat_sqrt_mean(const float* d2, int64_t m, double* partial, int n_partial, hipStream_t stream);
/* The same three over the pairs only: the n^2 distances are the n diagonal zeros and each pair twice,
 * so median and mean follow from the strict upper triangle (row < column). _upper computes the
 * tiles on and above the diagonal (each pair once, so the matrix read is exactly symmetric);
 * select2_upper takes ranks in [0, n(n-1)/2) over the strict upper triangle of the row-major d2;
 * sqrt_sum_upper sums sqrt(d2) over it (FP64 partials). n >= 2. */
int mocat_pairwise_sqdist_upper(const float* x, int64_t n, int64_t d, void* workspace, float* d2, hipStream_t stream);
int mocat_select2_upper(const float* d2, int64_t n, int64_t rank0, int64_t rank1, uint32_t* workspace,
                        hipStream_t stream);
int mocat_sqrt_sum_upper(const float* d2, int64_t n, double* partial, int n_partial, hipStream_t stream);

/* n_iter iterations of SVGD on a diagonal Gaussian mixture in one call (the loop of
 * mocat/src/sample.py:110-148 over SVGD.update, mocat/src/transport/svgd.py:122-146, with the
 * bandwidth adapted after every update as the reference's median/mean update override,
 * mocat/src/tests/test_transport.py:76-89): per iteration key <- split(key, n+2)[n], the Stein step
 * of mocat_svgd_step with the bandwidth read from *bandwidth, the mixture's potential and gradient
 * at the new particles (times temperature), then *bandwidth <- mean (bandwidth_mode 1) or median
 * (2) update, or unchanged (0). In/out: x, gsq, mom, potential, grad [n][d] / [n], key[2],
 * *bandwidth (device float). bandwidth_trace (optional) [n_iter]: the bandwidth after each
 * iteration. No host synchronisation. workspace: mocat_svgd_gmm_run_workspace_bytes. The median's
 * full digit passes (the first selection, and any whose bracket misses) run in one launch with a
 * bounded grid barrier; a barrier that times out (the MOCAT_SVGD_BARRIER_POLLS environment variable
 * lowers the wait limit; 0 forces it) makes that iteration's bandwidth NaN. */
int mocat_svgd_gmm_run_workspace_bytes(int64_t n, int64_t d, int js_count, int64_t* bytes);
int mocat_svgd_gmm_run(float* x, int64_t n, int64_t d, int K, const float* means, const float* lam, const float* logc,
                       float temperature, float step, float momentum, int bandwidth_mode, float* bandwidth,
                       int n_iter, int js_count, void* workspace, int64_t workspace_bytes, float* gsq, float* mom,
                       float* potential, float* grad, uint32_t* key, float* bandwidth_trace, hipStream_t stream);

/* ---- Kernel Stein discrepancy (replaces mocat/src/metrics.py:88-130, ksd with
 * ensemble_batchsize=None, for kernels.Gaussian (kernel 0) and kernels.IMQ (kernel 1) with mocat's
 * k_0 of mocat/src/metrics.py:116-124 and the kernel methods of mocat/src/kernels.py:83-111,
 * 164-215). part[js][i] = sum over the js-th share of 64-particle j tiles of
 * weights[j] * k_0(x_i, x_j) (FP64); weights may be NULL (all 1). The caller forms
 * sqrt(sum_i w_i sum_js part[js][i]) / sum(w). c and beta are used by the IMQ kernel only. */
int mocat_ksd_rows(const float* x, const float* grad_potential, const float* weights, int64_t n, int64_t d,
                   int kernel, float bandwidth, float c, float beta, int js_count, double* part,
                   hipStream_t stream);
/* ---- KSD over random minibatches (replaces mocat/src/metrics.py:101-104,126-128, ksd with
 * ensemble_batchsize = batch): inds[i][b] (int32, in [0, n)) are particle i's batch;
 * row_k[i] = sum_b weights[inds[i][b]] * k_0(x_i, x_inds[i][b]) and row_w[i] = sum_b
 * weights[inds[i][b]] (FP64, in b order; weights may be NULL). The caller forms
 * sqrt(sum_i w_i row_k[i] / (sum(w) row_w[i])). */
int mocat_ksd_batch_rows(const float* x, const float* grad_potential, const float* weights, int64_t n, int64_t d,
                         int64_t batch, const int32_t* inds, int kernel, float bandwidth, float c, float beta,
                         double* row_k, double* row_w, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MOCAT_AMD_H */
