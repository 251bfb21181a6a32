"""Optimal-proposal particle filter and ensemble Kalman filter, NumPy restatement (test
infrastructure only).

mocat's OptimalNonLinearGaussianParticleFilter and EnsembleKalmanFilter
(mocat/src/ssm/nonlinear_gaussian.py:134-281, 284-350) read a model's initial mean and covariance
square roots, likelihood matrix and covariance square roots, transition covariance and
transition_function. The restatement takes the linear transition f(x) = A x of
TimeHomogenousLinearGaussian (linear_gaussian.py:112-116), the model mocat's own examples run the
optimal filter on (examples/ssm/linear_gaussian_1D.py:55-68). Square roots follow reset_covariance
(mocat/src/utils.py:247-262): C_sqrt = cholesky(C), precision_sqrt = inv(cholesky(C)).

Key tree (run_particle_filter_for_marginals, mocat/src/ssm/filtering.py:255-324): as the bootstrap
filter, except that both filters draw all particles' noise from the FIRST key of each split:
  init  normal(split(sub, n)[0], (n, d))                         nonlinear_gaussian.py:198-209, 312-324
  step  optimal: normal(split(int_key, n)[0], (n, d))            :265-281
        EnKF:    (k1, k2) = split(split(int_key, n)[0]);         :326-350
                 normal(k1, (n, d)) transition noise, normal(k2, (n, dy)) observation perturbation
Initial log-weights are zero for both (:211-224, 312-324); the EnKF's stay zero (ESS = n, no
resampling at any ess_threshold <= 1).

Pinning (tests/test_oracle_nlg.py): the optimal filter's step satisfies, for any SPD Q and R,
  log q(x' | x, y) + w(x, y) = log f(x' | x) + log g(y | x')
with q, f, g independent Gaussian log-densities (a wrong gain, proposal covariance, weight precision
or normaliser breaks it); the startup matrices equal the conditioning formulas written in
information form; the filters' weighted / ensemble means and covariances converge to the Kalman
filter's (oracle.kalman); the optimal filter's log-likelihood estimate to log p(y_1:T | y_0); a 1-D
EnKF step equals its scalar closed form. For non-diagonal P0 or R mocat's initial conditioned
precision (init_prec_sqrt @ init_prec_sqrt + H^T R_prec_sqrt R_prec_sqrt^T H, :166-168) is not the
posterior precision; the restatement keeps mocat's expression.
"""
import numpy as np

from . import pf as opf
from . import rng

LOG2PI = np.log(2 * np.pi)


def _kalman_gain(cov_sqrt, H, lik_cov_sqrt, inv_mat=None):
    """mocat/src/utils.py:476-484."""
    if inv_mat is None:
        hc = H @ cov_sqrt
        inv_mat = np.linalg.inv(hc @ hc.T + lik_cov_sqrt @ lik_cov_sqrt.T)
    return cov_sqrt @ cov_sqrt.T @ H.T @ inv_mat


def _chol(M):
    """jnp.linalg.cholesky: the input is symmetrised, (M + M^T) / 2, before factoring."""
    return np.linalg.cholesky(0.5 * (M + M.T))


def startup(m0, P0, A, Q, H, R):
    """OptimalNonLinearGaussianParticleFilter.startup (nonlinear_gaussian.py:154-184) in FP64; the
    EnKF's startup (:297-310) is its first half (K0, C0)."""
    m0, P0, A, Q, H, R = (np.asarray(v, np.float64) for v in (m0, P0, A, Q, H, R))
    L0 = np.linalg.cholesky(P0)
    S0 = np.linalg.inv(L0)
    LQ = np.linalg.cholesky(Q)
    LR = np.linalg.cholesky(R)
    SR = np.linalg.inv(LR)
    K0 = _kalman_gain(L0, H, LR)
    hs = H.T @ SR
    icp = S0 @ S0 + hs @ hs.T
    C0 = np.linalg.inv(_chol(icp))
    hq = H @ LQ
    W = np.linalg.inv(hq @ hq.T + LR @ LR.T)
    K = _kalman_gain(LQ, H, LR, inv_mat=W)
    prop_cov = Q - K @ H @ Q
    S = _chol(prop_cov)
    WC = _chol(W)
    dy = H.shape[0]
    nlzw = (dy * LOG2PI - np.log(np.linalg.det(W))) / 2
    return dict(m0=m0, H=H, K0=K0, C0=C0, A=A, K=K, S=S, WC=WC, nlzw=nlzw, LQ=LQ, LR=LR,
                prop_cov=prop_cov, W=W, d=m0.shape[0], dy=dy)


def optimal_flat(p):
    """Kernel parameter layout [m0, H, K0, C0, A, K, S, WC, nlzw]."""
    return np.concatenate([p["m0"].ravel(), p["H"].ravel(), p["K0"].ravel(), p["C0"].ravel(), p["A"].ravel(),
                           p["K"].ravel(), p["S"].ravel(), p["WC"].ravel(), [p["nlzw"]]]).astype(np.float32)


def enkf_flat(p):
    """Kernel parameter layout [m0, H, K0, C0, A, LQ, LR]."""
    return np.concatenate([p["m0"].ravel(), p["H"].ravel(), p["K0"].ravel(), p["C0"].ravel(), p["A"].ravel(),
                           p["LQ"].ravel(), p["LR"].ravel()]).astype(np.float32)


def shared_normals(key, n, d):
    """normal(split(key, n)[0], (n, d)): every particle's noise from the first split key."""
    return rng.normal(rng.split(np.asarray(key, np.uint32), n)[0], (n, d)).astype(np.float64)


def initial(p, y0, z):
    """initial_sample (nonlinear_gaussian.py:198-209, 312-324): m0 + K0 (y0 - H m0) + z @ C0."""
    mc = p["m0"] + p["K0"] @ (np.asarray(y0, np.float64) - p["H"] @ p["m0"])
    return mc[None, :] + z @ p["C0"]


def optimal_step(p, xp, y, z):
    """propose_and_intermediate_weight_vectorised (nonlinear_gaussian.py:265-281): returns x_new and
    the log-weight increment -gaussian_potential(y, H A x, sqrt_prec = WC, det_prec = det W)."""
    mx = xp @ p["A"].T
    r = np.asarray(y, np.float64)[None, :] - mx @ p["H"].T
    x = mx + r @ p["K"].T + z @ p["S"].T
    v = r @ p["WC"]
    return x, -(0.5 * np.sum(v * v, axis=1) + p["nlzw"])


def run_optimal_pf(p, y, key, n, ess_threshold=0.5, resampling="systematic", ancestors=None,
                   resample_flags=None):
    """run_particle_filter_for_marginals with OptimalNonLinearGaussianParticleFilter. Returns
    dict(value (T, n, d), log_weight (T, n), ess (T,), log_likelihood, ancestors (T, n)); the
    log-likelihood estimate sum_t [lse(lw_t) - lse(lw_res_{t-1})] estimates log p(y_1:T | y_0) (the
    initial weights are zero). ancestors / resample_flags replay another run's resampling."""
    y = np.asarray(y, np.float64).reshape(len(y), -1)
    T, d = y.shape[0], p["d"]
    rkey, sub = rng.split(np.asarray(key, np.uint32), 2)
    x = initial(p, y[0], shared_normals(sub, n, d))
    lw = np.zeros(n)
    values, lws, ancs = [x], [lw], [np.arange(n, dtype=np.int32)]
    l0 = opf.lse(lw)
    esss = [np.exp(2 * l0 - opf.lse(2 * lw))]
    loglik = l0 - np.log(n)
    int_keys = rng.split(rkey, T - 1) if T > 1 else np.zeros((0, 2), np.uint32)
    for t in range(1, T):
        ik = int_keys[t - 1]
        res = esss[-1] < ess_threshold * n if resample_flags is None else bool(resample_flags[t])
        if res:
            if ancestors is not None:
                a = np.asarray(ancestors[t], np.int64)
            else:
                w = np.exp(lw - l0).astype(np.float32)
                a = opf.resample_indices(w, ik, resampling)
            lw_res, lres = np.zeros(n), np.log(n)
        else:
            a = np.arange(n, dtype=np.int32)
            lw_res, lres = lw, l0
        x, inc = optimal_step(p, x[a], y[t], shared_normals(ik, n, d))
        lw = lw_res + inc
        l0 = opf.lse(lw)
        loglik += l0 - lres
        esss.append(np.exp(2 * l0 - opf.lse(2 * lw)))
        values.append(x)
        lws.append(lw)
        ancs.append(a)
    return dict(value=np.stack(values), log_weight=np.stack(lws), ess=np.array(esss), log_likelihood=loglik,
                ancestors=np.stack(ancs))


def enkf_step(p, xp, y, z1, z2):
    """EnsembleKalmanFilter.propose_and_intermediate_weight_vectorised (nonlinear_gaussian.py:326-350):
    mx = A x + z1 LQ^T; spread = (mx - mean)^T / sqrt(n - 1); K = kalman_gain(spread, H, LR);
    x_new = mx + (y - (H mx + z2 LR^T)) K^T."""
    n = xp.shape[0]
    mx = xp @ p["A"].T + z1 @ p["LQ"].T
    spread = (mx - mx.mean(0)).T / np.sqrt(n - 1)
    K = _kalman_gain(spread, p["H"], p["LR"])
    y_prop = mx @ p["H"].T + z2 @ p["LR"].T
    return mx + (np.asarray(y, np.float64)[None, :] - y_prop) @ K.T, K


def run_enkf(p, y, key, n):
    """run_particle_filter_for_marginals with EnsembleKalmanFilter. Returns dict(value (T, n, d),
    gain (T - 1, d, dy))."""
    y = np.asarray(y, np.float64).reshape(len(y), -1)
    T, d, dy = y.shape[0], p["d"], p["dy"]
    rkey, sub = rng.split(np.asarray(key, np.uint32), 2)
    x = initial(p, y[0], shared_normals(sub, n, d))
    values, gains = [x], []
    int_keys = rng.split(rkey, T - 1) if T > 1 else np.zeros((0, 2), np.uint32)
    for t in range(1, T):
        k1, k2 = rng.split(rng.split(int_keys[t - 1], n)[0], 2)
        z1 = rng.normal(k1, (n, d)).astype(np.float64)
        z2 = rng.normal(k2, (n, dy)).astype(np.float64)
        x, K = enkf_step(p, x, y[t], z1, z2)
        values.append(x)
        gains.append(K)
    return dict(value=np.stack(values), gain=np.array(gains).reshape(T - 1, d, dy))
