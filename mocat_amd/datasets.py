"""Synthetic inputs of the benchmark workloads (seeded NumPy PCG64 generators, host side).

logistic_regression(seed, n, d): X ~ N(0, 1) (n x d), theta* ~ N(0, I_d),
    y ~ Bernoulli(sigmoid(X theta*)); returns float32 X, y and theta*.
"""
from __future__ import annotations

import numpy as np


def logistic_regression(seed: int, n: int, d: int):
    rs = np.random.default_rng(seed)
    X = rs.standard_normal((n, d)).astype(np.float32)
    theta_star = rs.standard_normal(d).astype(np.float32)
    z = X.astype(np.float64) @ theta_star.astype(np.float64)
    p = 0.5 * (1.0 + np.tanh(0.5 * z))
    y = (rs.uniform(size=n) < p).astype(np.float32)
    return X, y, theta_star


def lgssm_config(seed: int = 0, d: int = 4, T: int = 100):
    """Linear-Gaussian SSM of the particle-filter benchmark: key = PRNGKey(seed) split into
    (A, H, simulation) keys; A = 0.9 I + off-diagonal 0.1 * normal; H = 0.5 * normal (entries
    N(0, 1/4)); Q = 0.1 I; R = 0.5 I; x0 ~ N(0, I); observations at t = 0..T-1 simulated from the
    model. Returns (model, y (T, d), t (T,))."""
    import torch
    from . import random as mr
    from .ssm import TimeHomogenousLinearGaussian
    kA, kH, ks = mr.split(mr.PRNGKey(seed), 3)
    off = 0.1 * mr.normal(kA, (d, d)).double().cpu().numpy()
    A = 0.9 * np.eye(d) + off * (1 - np.eye(d))
    H = 0.5 * mr.normal(kH, (d, d)).double().cpu().numpy()
    model = TimeHomogenousLinearGaussian(np.zeros(d), np.eye(d), A, 0.1 * np.eye(d), H, 0.5 * np.eye(d))
    t = np.arange(T, dtype=np.float64)
    sim = model.simulate(t, ks)
    return model, sim.y, sim.t


def sv_batch_config(seed: int = 0, n_models: int = 256, n_min: int = 256, n_max: int = 65536,
                    t_min: int = 50, t_max: int = 1000):
    """Batched stochastic-volatility filters: per model phi ~ U[0.9, 0.99], sigma ~ U[0.1, 0.4],
    beta ~ U[0.5, 1], particles log-uniform in [n_min, n_max], lengths uniform in [t_min, t_max];
    observations simulated from each model (NumPy PCG64 stream of `seed`, x0 ~ N(0, 1)).
    Returns lists (models, ys, ns)."""
    from .ssm import StochasticVolatility
    rs = np.random.default_rng(seed)
    phi = rs.uniform(0.9, 0.99, n_models)
    sigma = rs.uniform(0.1, 0.4, n_models)
    beta = rs.uniform(0.5, 1.0, n_models)
    ns = np.exp(rs.uniform(np.log(n_min), np.log(n_max), n_models)).astype(np.int64)
    Ts = rs.integers(t_min, t_max + 1, n_models)
    models, ys = [], []
    for i in range(n_models):
        models.append(StochasticVolatility(1, 0.0, 1.0, phi[i], sigma[i] ** 2, beta[i] ** 2))
        x = np.empty(Ts[i])
        x[0] = rs.standard_normal()
        e = rs.standard_normal(Ts[i])
        for k in range(1, Ts[i]):
            x[k] = phi[i] * x[k - 1] + sigma[i] * e[k]
        ys.append((beta[i] * np.exp(0.5 * x) * rs.standard_normal(Ts[i])).astype(np.float32)[:, None])
    return models, ys, [int(v) for v in ns]
