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
