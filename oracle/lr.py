"""Bayesian logistic regression potential and gradient, FP64 NumPy (test infrastructure only).

The reference has no logistic-regression scenario; a user writes one as a `Scenario` subclass
(README.md:35-49 pattern) and mocat differentiates it with jax.grad (mocat/src/core.py:204-230).
Restated here in closed form:
  prior_potential(t)  = 0.5 * prec * |t|^2 + 0.5 * d * (log 2pi - log prec)   (gaussian_potential,
                        mocat/src/utils.py:49-105, with a scalar precision)
  lik_potential(t)    = sum_i softplus(x_i.t) - y_i * x_i.t
  grad lik            = X^T (sigmoid(X t) - y)
Pinned by central finite differences in FP64 and by the Bernoulli log-pmf written directly
(tests/test_oracle_lr.py).
"""
import numpy as np


def softplus(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def sigmoid(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def prior_potential(theta, prec):
    theta = np.asarray(theta, np.float64)
    d = theta.shape[-1]
    return 0.5 * prec * np.sum(theta * theta, axis=-1) + 0.5 * d * (np.log(2 * np.pi) - np.log(prec))


def lik_potential(theta, X, y):
    z = np.asarray(theta, np.float64) @ np.asarray(X, np.float64).T
    return np.sum(softplus(z) - np.asarray(y, np.float64) * z, axis=-1)


def lik_grad(theta, X, y):
    X = np.asarray(X, np.float64)
    z = np.asarray(theta, np.float64) @ X.T
    return (sigmoid(z) - np.asarray(y, np.float64)) @ X


def potential_and_grad(theta, X, y, prec, temperature=1.0):
    theta = np.asarray(theta, np.float64)
    pp = prior_potential(theta, prec)
    lp = lik_potential(theta, X, y)
    g = prec * theta + temperature * lik_grad(theta, X, y)
    return pp + temperature * lp, g, pp, lp


def make_dataset(seed, n, d, prior_var=10.0):
    """Synthetic data: X ~ N(0,1), theta* ~ N(0, I), y ~ Bernoulli(sigmoid(X theta*)) (NumPy PCG64)."""
    rs = np.random.default_rng(seed)
    X = rs.standard_normal((n, d)).astype(np.float32)
    theta_star = rs.standard_normal(d).astype(np.float32)
    p = sigmoid(X.astype(np.float64) @ theta_star.astype(np.float64))
    y = (rs.uniform(size=n) < p).astype(np.float32)
    return X, y, theta_star
