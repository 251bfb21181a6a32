"""State-space models with fused particle-filter kernels (mocat/src/ssm/ssm.py,
mocat/src/ssm/linear_gaussian/linear_gaussian.py, mocat/src/ssm/scenarios/stochastic_volatility.py).

Parameters are factored on the host in FP64 exactly as mocat's reset_covariance does
(mocat/src/utils.py:247-262: covariance_sqrt = cholesky(C), precision_sqrt = inv(cholesky(C)),
precision_det = 1 / det(C)) and flattened into the kernel's parameter layout.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import random as mr

LOG2PI = float(np.log(2 * np.pi))


def _np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy().astype(np.float64)
    return np.asarray(a, np.float64)


class StateSpaceModel:
    name = None
    dim = None
    dim_obs = None
    kernel_model = None  # 0 linear Gaussian, 1 stochastic volatility

    def __repr__(self):
        return f"mocat_amd.StateSpaceModel.{self.__class__.__name__}"

    def kernel_params(self) -> np.ndarray:
        raise NotImplementedError

    # -- per-particle model functions (mocat/src/ssm/ssm.py:47-81) on the filters' hook kernel; batched
    # over leading key / particle axes. A model written in Python has no kernel: HOOK_ERROR.
    def _hook_params(self):
        from .filtering import HOOK_ERROR
        if self.kernel_model not in (0, 1):
            raise NotImplementedError(HOOK_ERROR.format(what=type(self).__name__))
        return self.kernel_model, self.kernel_params()

    def initial_sample(self, t, random_key):
        """ssm.py:47-51: x_0 from one key ((2,) -> (d,)) or many ((..., 2) -> (..., d))."""
        from .filtering import _hook
        key = mr.as_key(random_key)
        lead = tuple(key.shape[:-1])
        n = int(np.prod(lead)) if lead else 1
        model, params = self._hook_params()
        x, _ = _hook(model, params, self, 0, None, None, key, n)
        return x.reshape(lead + (self.dim,))

    def transition_sample(self, x_previous, t_previous, t_new, random_key):
        """ssm.py:53-60: x_new given x_previous, one key per particle."""
        from .filtering import _hook, _step_scale
        key = mr.as_key(random_key)
        lead = tuple(key.shape[:-1])
        n = int(np.prod(lead)) if lead else 1
        model, params = self._hook_params()
        y0 = np.zeros(self.dim_obs, np.float32)  # the weight the hook also returns is not used here
        x, _ = _hook(model, params, self, 1, x_previous, y0, key, n, _step_scale(self, t_previous, t_new))
        return x.reshape(lead + (self.dim,))

    def likelihood_potential(self, x, y, t=None):
        """ssm.py:73-81: -log p(y | x) per particle (x (..., d) -> (...))."""
        from .filtering import _hook
        xt = torch.as_tensor(x, dtype=torch.float32)
        lead = tuple(xt.shape[:-1])
        n = int(np.prod(lead)) if lead else 1
        model, params = self._hook_params()
        _, lw = _hook(model, params, self, 2, xt, y, None, n)
        return lw.reshape(lead)

    def simulate(self, t_all, random_key):
        """StateSpaceModel.simulate (mocat/src/ssm/ssm.py:138-163): keys split(key, 2T); latent path
        from the first T keys, observations from the last T. Returns cdict(x, y, t)."""
        from ..core import cdict
        t_all = _np(t_all)
        T = len(t_all)
        keys = mr.split(random_key, 2 * T)
        eps_x = mr.normal(keys[:T].contiguous(), (self.dim,)).double().cpu().numpy()
        eps_y = mr.normal(keys[T:].contiguous(), (self.dim_obs,)).double().cpu().numpy()
        xs = [self._init_from_eps(eps_x[0])]
        for t in range(1, T):
            xs.append(self._transition_from_eps(xs[-1], eps_x[t], t_all[t] - t_all[t - 1]))
        xs = np.stack(xs)
        ys = np.stack([self._observe_from_eps(xs[t], eps_y[t]) for t in range(T)])
        dev = mr.default_device()
        return cdict(x=torch.tensor(xs, dtype=torch.float32, device=dev),
                     y=torch.tensor(ys, dtype=torch.float32, device=dev),
                     t=torch.tensor(t_all, dtype=torch.float32, device=dev), name=f"{self.name} simulation")


class TimeHomogenousLinearGaussian(StateSpaceModel):
    """x_t ~ N(A x_{t-1}, Q), y_t ~ N(H x_t, R), x_0 ~ N(m0, P0); 2-D covariance matrices."""
    name = "Time-homogenous Linear Gaussian"
    kernel_model = 0

    def __init__(self, initial_mean=None, initial_covariance=None, transition_matrix=None,
                 transition_covariance=None, likelihood_matrix=None, likelihood_covariance=None, name=None):
        if name is not None:
            self.name = name
        dim = None
        for v in (initial_mean, initial_covariance, transition_covariance, likelihood_matrix):
            if v is not None:
                dim = _np(v).shape[-1]
                break
        if dim is None:
            raise AttributeError("Could not find dimension for TimeHomogenousLinearGaussian")
        self.dim = int(dim)
        self.initial_mean = np.zeros(dim) if initial_mean is None else _np(initial_mean)
        self.initial_covariance = np.eye(dim) if initial_covariance is None else _np(initial_covariance)
        self.transition_matrix = np.eye(dim) if transition_matrix is None else _np(transition_matrix)
        self.transition_covariance = np.eye(dim) if transition_covariance is None else _np(transition_covariance)
        self.likelihood_matrix = np.eye(dim) if likelihood_matrix is None else _np(likelihood_matrix)
        self.dim_obs = int(self.likelihood_matrix.shape[0])
        self.likelihood_covariance = np.eye(self.dim_obs) if likelihood_covariance is None else \
            _np(likelihood_covariance)
        for c in (self.initial_covariance, self.transition_covariance, self.likelihood_covariance):
            if c.ndim != 2:
                raise ValueError("covariances must be 2-D matrices")
        self.initial_covariance_sqrt = np.linalg.cholesky(self.initial_covariance)
        self.transition_covariance_sqrt = np.linalg.cholesky(self.transition_covariance)
        self.likelihood_covariance_sqrt = np.linalg.cholesky(self.likelihood_covariance)
        self.likelihood_precision_sqrt = np.linalg.inv(self.likelihood_covariance_sqrt)
        self.likelihood_precision_det = 1.0 / np.linalg.det(self.likelihood_covariance)

    def kernel_params(self):
        nlz = (self.dim_obs * LOG2PI - np.log(self.likelihood_precision_det)) / 2
        return np.concatenate([self.initial_mean.ravel(), self.initial_covariance_sqrt.ravel(),
                               self.transition_matrix.ravel(), self.transition_covariance_sqrt.ravel(),
                               self.likelihood_matrix.ravel(), self.likelihood_precision_sqrt.ravel(),
                               [nlz]]).astype(np.float32)

    def transition_params(self):
        """Backward-simulation parameters: A, the transition precision square root inv(cholesky(Q))
        and the normaliser (d log 2 pi - log det Q^-1) / 2 of mocat's gaussian_potential
        (linear_gaussian.py:73-84, utils.py:49-105)."""
        S = np.linalg.inv(self.transition_covariance_sqrt)
        nlz = (self.dim * LOG2PI - np.log(1.0 / np.linalg.det(self.transition_covariance))) / 2
        return np.concatenate([self.transition_matrix.ravel(), S.ravel(), [nlz]]).astype(np.float32)

    def _init_from_eps(self, e):
        return self.initial_covariance_sqrt @ e + self.initial_mean

    def _transition_from_eps(self, x, e, dt):
        return self.transition_matrix @ x + self.transition_covariance_sqrt @ e

    def _observe_from_eps(self, x, e):
        return self.likelihood_matrix @ x + self.likelihood_covariance_sqrt @ e


class StochasticVolatility(StateSpaceModel):
    """Stochastic volatility (mocat/src/ssm/scenarios/stochastic_volatility.py:15-108), dim = dim_obs <= 4:
    x_t = A x_{t-1} + chol(Q) eps sqrt(t_t - t_{t-1}), y_t,k = sqrt(R_k) exp(x_t,k / 2) eta_k. Arguments as
    mocat's: initial_mean, initial_covariance, transition_matrix (A), transition_covariance (Q),
    likelihood_covariance_diag (R; a matrix's diagonal is taken). In 1-D they are the scalars phi = A,
    sigma^2 = Q, beta^2 = R (attributes phi, sigma, beta; the log-likelihood gradient and backward
    simulation are compiled for this case). The transition noise scales with sqrt(t_new - t_previous)
    (unit steps when no times are given)."""
    name = "Stochastic Volatility"
    kernel_model = 1

    def __init__(self, dim: int = 1, initial_mean=None, initial_covariance=None, transition_matrix=None,
                 transition_covariance=None, likelihood_covariance_diag=None, name=None):
        d = int(dim)
        if not 1 <= d <= 4:
            raise NotImplementedError("the fused kernels cover the stochastic-volatility model for dim 1 to 4")
        if name is not None:
            self.name = name
        self.dim = self.dim_obs = d
        if d == 1:
            f = lambda v, default: default if v is None else float(np.asarray(_np(v)).ravel()[0])
            self.initial_mean = f(initial_mean, 0.0)
            self.initial_covariance = f(initial_covariance, 1.0)
            self.phi = f(transition_matrix, 1.0)
            self.sigma = float(np.sqrt(f(transition_covariance, 1.0)))
            self.beta = float(np.sqrt(f(likelihood_covariance_diag, 1.0)))
            return
        mat = lambda v, default: default if v is None else _np(v).reshape(d, d)
        self.initial_mean = np.zeros(d) if initial_mean is None else _np(initial_mean).reshape(d)
        self.initial_covariance = mat(initial_covariance, np.eye(d))
        self.transition_matrix = mat(transition_matrix, np.eye(d))
        self.transition_covariance = mat(transition_covariance, np.eye(d))
        r = np.ones(d) if likelihood_covariance_diag is None else _np(likelihood_covariance_diag)
        self.likelihood_covariance_diag = np.diag(r).copy() if r.ndim == 2 else r.reshape(d)
        # mocat's reset_covariance (utils.py:247-262): Cholesky factors in FP64
        self.initial_covariance_sqrt = np.linalg.cholesky(self.initial_covariance)
        self.transition_covariance_sqrt = np.linalg.cholesky(self.transition_covariance)
        self.likelihood_precision_diag_sqrt = 1.0 / np.sqrt(self.likelihood_covariance_diag)

    def kernel_params(self):
        if self.dim == 1:
            return np.array([self.initial_mean, np.sqrt(self.initial_covariance), self.phi, self.sigma,
                             1.0 / self.beta], np.float32)
        return np.concatenate([self.initial_mean, self.initial_covariance_sqrt.ravel(), self.transition_matrix.ravel(),
                               self.transition_covariance_sqrt.ravel(),
                               self.likelihood_precision_diag_sqrt]).astype(np.float32)

    def transition_params(self):
        """Backward-simulation parameters for unit steps: phi, 1/sigma and the normaliser
        (log 2 pi - log sigma^-2) / 2 (stochastic_volatility.py:78-84, utils.py:49-105)."""
        if self.dim != 1:
            raise NotImplementedError("backward simulation is compiled for the 1-D stochastic-volatility model")
        return np.array([self.phi, 1.0 / self.sigma, (LOG2PI + 2.0 * np.log(self.sigma)) / 2], np.float32)

    def _init_from_eps(self, e):
        if self.dim == 1:
            return np.sqrt(self.initial_covariance) * e + self.initial_mean
        return self.initial_covariance_sqrt @ e + self.initial_mean

    def _transition_from_eps(self, x, e, dt):
        if self.dim == 1:
            return x * self.phi + e * self.sigma * np.sqrt(dt)
        return self.transition_matrix @ x + (self.transition_covariance_sqrt @ e) * np.sqrt(dt)

    def _observe_from_eps(self, x, e):
        if self.dim == 1:
            return self.beta * np.exp(0.5 * x) * e
        return np.sqrt(self.likelihood_covariance_diag) * np.exp(0.5 * x) * e
