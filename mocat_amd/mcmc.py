"""MCMC samplers mirroring mocat/src/mcmc/{sampler,metropolis,standard_mcmc}.py, executed by the
fused HIP iteration kernel (libmocat_amd: mocat_hmc_lr_iterate).

`Underdamped(stepsize, leapfrog_steps, friction)` is underdamped Langevin with a Metropolis
correction; friction = inf gives HMC, and additionally leapfrog_steps = 1 gives MALA
(mocat/src/mcmc/standard_mcmc.py:72-153). Corrections: `Uncorrected`, `Metropolis`,
`RMMetropolis` (log Robbins-Monro stepsize adaptation toward `tuning.target`,
mocat/src/mcmc/metropolis.py:73-161).

One call of the kernel performs whole iterations (always -> leapfrog -> accept/reject -> adapt)
for tiles of 32 chains; the data contraction inside each leapfrog step runs on MFMA (bf16x3) or
vector FP32, chosen by the scenario's `precision`.
"""
from __future__ import annotations

import math
import time
from inspect import isclass
from typing import Union

import numpy as np
import torch

from . import _lib
from . import random as mr
from .core import Scenario, cdict, is_implemented
from .sample import Sampler

_lib.register("mocat_hmc_lr_iterate",
              [_lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_i32, _lib.c_f32, _lib.c_f32,
               _lib.c_i64, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr,
               _lib.c_ptr, _lib.c_ptr, _lib.c_i64, _lib.c_i32, _lib.c_i32, _lib.c_f32, _lib.c_i32, _lib.c_f32,
               _lib.c_f32, _lib.c_f32, _lib.c_f64, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr,
               _lib.c_ptr])

DEFAULT_STACK = ("value", "potential", "alpha", "stepsize", "momenta")


class Correction:
    def __init__(self, **kwargs):
        self.__dict__.update(kwargs)

    def __repr__(self):
        return f"mocat_amd.Correction.{self.__class__.__name__}"

    kernel_mode = None


class Uncorrected(Correction):
    kernel_mode = 0


class Metropolis(Correction):
    kernel_mode = 1


class RMMetropolis(Correction):
    kernel_mode = 2

    def __init__(self, super_correction=None, adapt_cut_off: float = math.inf, rm_stepsize_scale: float = 1.0,
                 rm_stepsize_neg_exponent: float = 2 / 3, log_update: bool = True):
        super().__init__()
        sc = Metropolis() if super_correction is None else super_correction
        self.super_correction = sc() if isclass(sc) else sc
        if not isinstance(self.super_correction, Metropolis):
            raise NotImplementedError("RMMetropolis wraps Metropolis in the fused kernel")
        if not log_update:
            raise NotImplementedError("the fused kernel implements the log Robbins-Monro update")
        self.adapt_cut_off = adapt_cut_off
        self.rm_stepsize_scale = rm_stepsize_scale
        self.rm_stepsize_neg_exponent = rm_stepsize_neg_exponent
        self.log_update = log_update


def check_correction(correction) -> Correction:
    if correction is None:
        return Uncorrected()
    if isclass(correction) and issubclass(correction, Correction):
        return correction()
    if not isinstance(correction, Correction):
        raise TypeError("Correction must be of type mocat_amd.Correction")
    return correction


class MCMCSampler(Sampler):
    correction = None

    def __init__(self, **kwargs):
        if "tuning" not in self.__dict__:
            self.tuning = cdict(parameter="stepsize", target=None, metric="alpha", monotonicity="decreasing")
        super().__init__(**kwargs)


class Underdamped(MCMCSampler):
    name = "Underdamped"
    correction = Metropolis

    def __init__(self, stepsize: float = None, leapfrog_steps: int = 1, friction: float = math.inf):
        super().__init__()
        self.parameters.stepsize = stepsize
        self.parameters.leapfrog_steps = int(leapfrog_steps)
        self.parameters.friction = friction
        self.tuning.target = 0.651

    # -------------------------------------------------------------------------------------------
    def startup(self, scenario: Scenario, n: int, initial_state, initial_extra: cdict, n_chains=None, **kwargs):
        """Batched MCMCSampler.startup + correction.startup + Underdamped.startup
        (mocat/src/mcmc/sampler.py:32-61, metropolis.py:35-45,91-116, standard_mcmc.py:86-105)."""
        if "correction" in kwargs:
            self.correction = kwargs.pop("correction")
        self.correction = check_correction(self.correction)
        self.max_iter = n - 1
        d = scenario.dim
        key = initial_extra.random_key
        keys = mr.split(key, n_chains) if n_chains else key.reshape(1, 2)
        C = keys.shape[0]
        if initial_state is None:
            if is_implemented(scenario.prior_sample):
                s = mr.split(keys, 2)
                keys = s[:, 0].contiguous()
                vals = scenario.prior_sample(s[:, 1].contiguous())
            else:
                vals = torch.zeros((C, d), dtype=torch.float32, device=keys.device)
            initial_state = cdict(value=vals)
        else:
            initial_state = initial_state.copy()
            initial_state.value = torch.as_tensor(initial_state.value, dtype=torch.float32,
                                                  device=keys.device).reshape(C, d).contiguous()
        initial_state, initial_extra = super().startup(scenario, n, initial_state, initial_extra, **kwargs)
        params = initial_extra.parameters
        if isinstance(self.correction, RMMetropolis):
            if params.stepsize is None:
                params.stepsize = 2.38 ** 2 / d
            self.tuning.monotonicity = -1
        if params.stepsize is None:
            raise ValueError("stepsize must be given unless the correction is RMMetropolis")
        initial_state.stepsize = torch.full((C,), float(params.stepsize), dtype=torch.float32, device=keys.device)
        initial_state.alpha = torch.ones(C, dtype=torch.float32, device=keys.device)
        # (key, prior_key, lik_key) = split(key, 3); the fused scenarios take no potential keys
        keys = mr.split(keys, 3)[:, 0].contiguous()
        T = float(scenario.temperature)
        pp, lp, gp, gl = scenario.evaluate(initial_state.value, T)
        initial_state.prior_potential = pp
        initial_state.likelihood_potential = lp
        initial_state.potential = pp + T * lp
        initial_state.grad_potential = (gp + T * gl).contiguous()
        if not hasattr(initial_state, "momenta") or initial_state.momenta.shape[-1] != d:
            initial_state.momenta = torch.zeros((C, d), dtype=torch.float32, device=keys.device)
        else:
            initial_state.momenta = torch.as_tensor(initial_state.momenta, dtype=torch.float32,
                                                    device=keys.device).reshape(C, d).contiguous()
        initial_extra.random_key = keys
        return initial_state, initial_extra

    def _correction_args(self):
        c = self.correction
        if isinstance(c, RMMetropolis):
            return 2, float(self.tuning.target), float(c.rm_stepsize_scale), float(c.rm_stepsize_neg_exponent), \
                float(c.adapt_cut_off)
        return c.kernel_mode, float(self.tuning.target or 0.0), 1.0, 2 / 3, math.inf

    def iterate(self, scenario, state: cdict, extra: cdict, n_iter: int, traces: dict = None):
        """Advance the batched state in place by n_iter iterations (one kernel launch)."""
        if not hasattr(scenario, "kernel_args"):
            raise NotImplementedError(f"{scenario.__class__.__name__} has no fused HIP kernel")
        traces = traces or {}
        X, y, packed, nrows, d, prec, pprec = scenario.kernel_args()
        C = state.value.shape[0]
        mode, target, rm_scale, rm_exp, cut = self._correction_args()
        p = extra.parameters
        _lib.call("mocat_hmc_lr_iterate", X, y, packed, nrows, d, prec, pprec, float(scenario.temperature), C,
                  _lib.ptr(state.value), _lib.ptr(state.momenta), _lib.ptr(state.grad_potential),
                  _lib.ptr(state.potential), _lib.ptr(state.prior_potential), _lib.ptr(state.likelihood_potential),
                  _lib.ptr(state.stepsize), _lib.ptr(state.alpha), _lib.ptr(extra.random_key), int(extra.iter),
                  int(n_iter), int(p.leapfrog_steps), float(p.friction), mode, target, rm_scale, rm_exp, cut,
                  _lib.ptr(traces.get("value")), _lib.ptr(traces.get("potential")), _lib.ptr(traces.get("alpha")),
                  _lib.ptr(traces.get("stepsize")), _lib.ptr(traces.get("momenta")), _lib.stream_handle())
        extra.iter += int(n_iter)
        return state, extra

    def update(self, scenario, state: cdict, extra: cdict):
        """One MCMCSampler.update (mocat/src/mcmc/sampler.py:92-111) for every chain in `state`."""
        return self.iterate(scenario, state, extra, 1)

    def run_fused(self, scenario, n, initial_state, initial_extra, n_chains=None, stack=DEFAULT_STACK,
                  iterations_per_launch: int = None, **kwargs):
        state, extra = self.startup(scenario, n, initial_state, initial_extra, n_chains=n_chains, **kwargs)
        summary = self.summary(scenario, state, extra)
        C, d = state.value.shape
        n_it = n - 1
        dev = state.value.device
        shapes = {"value": (n_it, C, d), "momenta": (n_it, C, d), "potential": (n_it, C), "alpha": (n_it, C),
                  "stepsize": (n_it, C)}
        traces = {k: torch.empty(shapes[k], dtype=torch.float32, device=dev) for k in stack}
        init = state.copy()
        init_snap = {k: getattr(state, k).clone() for k in
                     ("value", "momenta", "potential", "alpha", "stepsize", "prior_potential",
                      "likelihood_potential", "grad_potential")}
        per = n_it if not iterations_per_launch else int(iterations_per_launch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < n_it:
            m = min(per, n_it - done)
            view = {k: v[done:done + m] for k, v in traces.items()}
            self.iterate(scenario, state, extra, m, view)
            done += m
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        for k, v in init_snap.items():
            setattr(init, k, v)
        chain = cdict()
        for k in stack:
            chain.__dict__[k] = torch.cat([getattr(init, k)[None], traces[k]], dim=0)
        chain.final_state = state
        chain.final_extra = extra
        if n_chains is None:
            for k in stack:
                chain.__dict__[k] = chain.__dict__[k][:, 0]
        return chain, elapsed, summary
