"""Scenarios with fused GPU potential + gradient kernels.

LogisticRegression: Bayesian logistic regression with a N(0, prior_variance I) prior. The reference
has no such scenario: it is the user Scenario pattern of README.md:35-49 (prior + likelihood
potentials, gradients by jax.grad via mocat/src/core.py:204-230), here with a hand-written adjoint
(grad = X^T (sigmoid(X theta) - y) + theta / prior_variance) evaluated by libmocat_amd.
"""
from __future__ import annotations

import math

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from . import _lib
from . import random as mr
from .core import Scenario

PRECISIONS = {"fp32": 0, "bf16x3": 1, "fp64": 3}
ENGINE_GAUSSIAN = 2

_lib.register("mocat_gmm_potential_grad", [_lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_i32, _lib.c_ptr, _lib.c_ptr,
                                           _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr])
_lib.register("mocat_gmm_full_potential_grad", [_lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_i32, _lib.c_ptr,
                                                _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr])
_lib.register("mocat_lr_packed_bytes", [_lib.c_i64, _lib.c_i64, _lib.c_ptr])
_lib.register("mocat_lr_pack", [_lib.c_ptr, _lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_ptr, _lib.c_ptr])
_lib.register("mocat_lr_potential_grad",
              [_lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_i32, _lib.c_f32, _lib.c_f32,
               _lib.c_ptr, _lib.c_i64, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr])


class FusedPotential(torch.autograd.Function):
    """A fused scenario's potential as a differentiable function of x (the torch counterpart of the
    jax.custom_vjp that mocat's jax.grad would otherwise differentiate through,
    mocat/src/core.py:204-230). The forward pass runs the scenario's kernel, which evaluates the
    potential and its hand-written adjoint together; the backward pass scales that adjoint by the
    incoming gradient. `part` selects "prior", "likelihood" or "tempered" (prior + temperature * lik).
    Second derivatives are not provided (the kernels have no second-order adjoint)."""

    @staticmethod
    def forward(ctx, x, scenario, temperature, part):
        pp, lp, gp, gl = scenario.evaluate(x.detach(), temperature)
        if part == "prior":
            val, g = pp, gp
        elif part == "likelihood":
            val, g = lp, gl
        elif part == "tempered":
            val, g = pp + temperature * lp, gp + temperature * gl
        else:
            raise ValueError(f"unknown potential part {part!r}")
        ctx.save_for_backward(g)
        ctx.x_device = x.device
        return val

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_val):
        (g,) = ctx.saved_tensors
        return (grad_val.to(g.device).unsqueeze(-1) * g).to(ctx.x_device), None, None, None


def _dev_tensor(a, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        t = a.detach().to(dtype)
    else:
        t = torch.as_tensor(np.asarray(a), dtype=dtype)
    if not t.is_cuda:
        t = t.to(mr.default_device())
    return t.contiguous()


class LogisticRegression(Scenario):
    """Bayesian logistic regression: y_i ~ Bernoulli(sigmoid(x_i . theta)), theta ~ N(0, v I).

    precision: "bf16x3" (MFMA, hi/lo split, FP32 accumulation), "fp32" (vector FMAs) or "fp64" (vector
    FMAs in double with f32 inputs and outputs: the reference the other two are measured against).
    Supports d <= 128.
    """
    name = "Logistic Regression"
    _fused_kernel = True

    def __init__(self, X, y, prior_variance: float = 10.0, precision: str = "bf16x3", name: str = None):
        super().__init__(name=name)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        self.X = _dev_tensor(X)
        self.y = _dev_tensor(y).reshape(-1)
        if self.X.dim() != 2 or self.X.shape[0] != self.y.shape[0]:
            raise ValueError("X must be (n, d) and y (n,)")
        self.n, self.dim = int(self.X.shape[0]), int(self.X.shape[1])
        if self.dim > 128:
            raise ValueError("LogisticRegression kernels support d <= 128")
        self.prior_variance = float(prior_variance)
        self.prior_precision = 1.0 / self.prior_variance
        self.precision = precision
        self.packed = None
        if precision == "bf16x3":
            nbytes = _lib.ctypes.c_int64(0)
            _lib.call("mocat_lr_packed_bytes", self.n, self.dim, _lib.ctypes.byref(nbytes))
            self.packed = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.X.device)
            _lib.call("mocat_lr_pack", _lib.ptr(self.X), _lib.ptr(self.y), self.n, self.dim, _lib.ptr(self.packed),
                      _lib.stream_handle())

    # kernel arguments shared with the fused samplers
    def kernel_args(self):
        return (_lib.ptr(self.X), _lib.ptr(self.y), _lib.ptr(self.packed), self.n, self.dim,
                PRECISIONS[self.precision], float(self.prior_precision))

    def evaluate(self, x: torch.Tensor, temperature: float = 1.0):
        """(prior_potential, likelihood_potential, grad_prior, grad_likelihood) for x (..., d)."""
        x = _dev_tensor(x)
        if x.shape[-1] != self.dim:
            raise ValueError(f"expected trailing dimension {self.dim}, got {tuple(x.shape)}")
        batch = x.shape[:-1]
        C = int(math.prod(batch)) if len(batch) else 1
        pp = torch.empty(batch, dtype=torch.float32, device=x.device)
        lp = torch.empty_like(pp)
        gp = torch.empty_like(x)
        gl = torch.empty_like(x)
        X, y, packed, n, d, prec, pprec = self.kernel_args()
        _lib.call("mocat_lr_potential_grad", X, y, packed, n, d, prec, pprec, float(temperature), _lib.ptr(x), C,
                  _lib.ptr(pp), _lib.ptr(lp), _lib.ptr(gp), _lib.ptr(gl), _lib.stream_handle())
        return pp, lp, gp, gl

    # differentiable potentials: torch.autograd reaches the kernel's adjoint through FusedPotential
    def prior_potential(self, x, random_key=None):
        return FusedPotential.apply(x, self, 1.0, "prior")

    def likelihood_potential(self, x, random_key=None):
        return FusedPotential.apply(x, self, 1.0, "likelihood")

    def tempered_potential(self, x, temperature, random_key=None):
        return FusedPotential.apply(x, self, float(temperature), "tempered")

    def prior_potential_and_grad(self, x, random_key=None):
        pp, _, gp, _ = self.evaluate(x)
        return pp, gp

    def likelihood_potential_and_grad(self, x, random_key=None):
        _, lp, _, gl = self.evaluate(x)
        return lp, gl

    def tempered_potential_and_grad(self, x, temperature, random_key=None):
        pp, lp, gp, gl = self.evaluate(x, temperature)
        return pp + temperature * lp, gp + temperature * gl

    def prior_sample(self, random_key):
        """sqrt(v) * normal(key, (d,)), batched over keys (..., 2)."""
        return math.sqrt(self.prior_variance) * mr.normal(random_key, (self.dim,))


class DiagonalGaussian(LogisticRegression):
    """Gaussian likelihood with diagonal covariance and a N(0, prior_variance I) prior.

    likelihood_potential(x) = 0.5 * sum_k (x_k - mean_k)^2 / covariance_k (no normalising constant, as
    mocat's toy Gaussian, mocat/src/scenarios/toy_examples.py:19-51); prior_variance = inf gives mocat's
    improper flat prior (potential 0). Evaluated by the same fused kernels (elementwise engine).
    """
    name = "Diagonal Gaussian"

    def __init__(self, mean, covariance, prior_variance: float = 10.0, name: str = None):
        Scenario.__init__(self, name=name)
        self.mean_t = _dev_tensor(mean).reshape(-1)
        cov = _dev_tensor(covariance).reshape(-1)
        if cov.shape != self.mean_t.shape:
            raise ValueError("mean and covariance (diagonal) must have the same length")
        self.lam = (1.0 / cov).contiguous()
        self.dim = int(self.mean_t.shape[0])
        if self.dim > 128:
            raise ValueError("the fused kernels support d <= 128")
        self.n = 1
        self.prior_variance = float(prior_variance)
        self.prior_precision = 0.0 if math.isinf(self.prior_variance) else 1.0 / self.prior_variance
        self.precision = "fp32"
        self.packed = None
        if math.isinf(self.prior_variance):
            # flat prior: no prior_sample (samplers then start from N(0, I), as mocat does)
            self.prior_sample = Scenario.prior_sample.__get__(self)

    def kernel_args(self):
        return (_lib.ptr(self.mean_t), _lib.ptr(self.lam), _lib.ptr(None), 1, self.dim, ENGINE_GAUSSIAN,
                float(self.prior_precision))

    def prior_sample(self, random_key):
        return math.sqrt(self.prior_variance) * mr.normal(random_key, (self.dim,))


class GaussianMixture(Scenario):
    """Gaussian mixture (mocat/src/scenarios/toy_examples.py:54-116).

    potential(x) = -log sum_k w_k c_k exp(-0.5 (x - mu_k)^T S_k (x - mu_k)), evaluated in log-sum-exp
    form. det_convention="textbook": S_k = P_k = inv(Sigma_k) and c_k = sqrt(det P_k) (2 pi)^(-d/2) (a
    normalised mixture); "mocat": the reference's component potential as written (toy_examples.py:95-103),
    c_k = det(P_k) (2 pi)^(-d/2) and the quadratic form |L_k (x - mu_k)|^2 of L_k = cholesky(P_k), i.e.
    S_k = L_k^T L_k (equal to P_k for diagonal covariances). Diagonal covariances (given as [K][d]
    variances or diagonal matrices) run on mocat_gmm_potential_grad and the one-call SVGD run; full
    covariances on mocat_gmm_full_potential_grad.
    No prior: prior_potential = 0 (mocat warns and assumes a uniform prior, core.py:183-189).
    """
    name = "Gaussian Mixture"
    _fused_kernel = True

    def __init__(self, means, covariances=None, weights=None, det_convention: str = "textbook", name: str = None):
        super().__init__(name=name)
        mu = np.asarray(means.detach().cpu() if isinstance(means, torch.Tensor) else means, np.float64)
        if mu.ndim == 1:
            mu = mu[:, None]
        K, d = mu.shape
        if K > 8:
            raise ValueError("at most 8 components")
        full = None
        if covariances is None:
            var = np.ones((K, d))
        else:
            c = np.asarray(covariances.detach().cpu() if isinstance(covariances, torch.Tensor) else covariances,
                           np.float64)
            if c.ndim == 2 and c.shape == (d, d):  # one covariance matrix for every component (as mocat)
                c = np.repeat(c[None], K, 0)
            if c.ndim == 3:
                var = np.einsum("kii->ki", c)
                if not np.allclose(c, np.einsum("ki,ij->kij", var, np.eye(d)), rtol=0.0, atol=0.0):
                    full = c
            else:  # [K][d] (or broadcastable) variances
                var = np.broadcast_to(c, (K, d)).astype(np.float64)
        w = np.ones(K) / K if weights is None else np.asarray(weights, np.float64)
        dev = mr.default_device()
        self.dim, self.n_components = int(d), int(K)
        self.means = torch.tensor(mu, dtype=torch.float32, device=dev).contiguous()
        self.weights = w
        self.full = full is not None
        if self.full:
            # the component matrices in FP64 on the host (mocat's __setattr__, toy_examples.py:73-88):
            # P = inv(Sigma), L = cholesky(P) (input symmetrised, as jnp.linalg.cholesky reads one triangle)
            P = np.linalg.inv(full)
            L = np.linalg.cholesky(0.5 * (P + np.transpose(P, (0, 2, 1))))
            detP = np.linalg.det(P)
            if det_convention == "mocat":
                S = np.einsum("kji,kjl->kil", L, L)
                logc = np.log(w * detP) - 0.5 * d * np.log(2 * np.pi)
            else:
                S = 0.5 * (P + np.transpose(P, (0, 2, 1)))
                logc = np.log(w) + 0.5 * np.log(detP) - 0.5 * d * np.log(2 * np.pi)
            self.S = torch.tensor(S, dtype=torch.float32, device=dev).contiguous()
            self.lam = None
            self.covariances = full
        else:
            lam = 1.0 / var
            factor = {"textbook": 0.5, "mocat": 1.0}[det_convention]
            logc = np.log(w) + factor * np.log(lam).sum(1) - 0.5 * d * np.log(2 * np.pi)
            self.lam = torch.tensor(lam, dtype=torch.float32, device=dev).contiguous()
            self.S = None
        self.logc = torch.tensor(logc, dtype=torch.float32, device=dev).contiguous()
        self.variances = var

    def evaluate(self, x, temperature: float = 1.0):
        x = _dev_tensor(x)
        batch = x.shape[:-1]
        n = int(math.prod(batch)) if len(batch) else 1
        pot = torch.empty(batch, dtype=torch.float32, device=x.device)
        grad = torch.empty_like(x)
        if self.full:
            _lib.call("mocat_gmm_full_potential_grad", _lib.ptr(x), n, self.dim, self.n_components,
                      _lib.ptr(self.means), _lib.ptr(self.S), _lib.ptr(self.logc), _lib.ptr(pot), _lib.ptr(grad),
                      _lib.stream_handle())
        else:
            _lib.call("mocat_gmm_potential_grad", _lib.ptr(x), n, self.dim, self.n_components, _lib.ptr(self.means),
                      _lib.ptr(self.lam), _lib.ptr(self.logc), _lib.ptr(pot), _lib.ptr(grad), _lib.stream_handle())
        zeros = torch.zeros_like(pot)
        return zeros, pot, torch.zeros_like(grad), grad

    def prior_potential(self, x, random_key=None):
        return torch.zeros(x.shape[:-1], dtype=torch.float32, device=x.device)

    def likelihood_potential(self, x, random_key=None):
        return FusedPotential.apply(x, self, 1.0, "likelihood")

    def tempered_potential(self, x, temperature, random_key=None):
        return FusedPotential.apply(x, self, float(temperature), "tempered")

    def prior_potential_and_grad(self, x, random_key=None):
        return self.prior_potential(x), torch.zeros_like(x)

    def likelihood_potential_and_grad(self, x, random_key=None):
        _, lp, _, gl = self.evaluate(x)
        return lp, gl

    def potential_and_grad(self, x, random_key=None):
        _, lp, _, gl = self.evaluate(x)
        return self.temperature * lp, self.temperature * gl
