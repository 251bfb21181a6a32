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

_lib.register("mocat_hmc_begin", [_lib.c_i64, _lib.c_i32, _lib.c_i32, _lib.c_f32, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr,
                                   _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr])
_lib.register("mocat_hmc_kick", [_lib.c_i64, _lib.c_i32, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr, _lib.c_i32,
                                  _lib.c_ptr])
_lib.register("mocat_hmc_correct",
              [_lib.c_i64, _lib.c_i32, _lib.c_i64, _lib.c_i32, _lib.c_f32, _lib.c_f32, _lib.c_f32, _lib.c_f64]
              + [_lib.c_ptr] * 17)

DEFAULT_STACK = ("value", "potential", "alpha", "stepsize", "momenta")


def evaluate(scenario, x, temperature, random_key=None):
    """(prior_potential, likelihood_potential, grad_prior, grad_likelihood) at x (C, d): the fused
    kernel of a compiled scenario, or a user scenario's torch potentials (their keys as mocat passes
    them: `random_key` [C][2] per chain, or None)."""
    if getattr(scenario, "_fused_kernel", False):
        return scenario.evaluate(x, temperature)
    return scenario.evaluate(x, temperature, random_key)


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
        # (key, prior_key, lik_key) = split(key, 3); mocat passes prior_key to both potentials (the fused
        # scenarios take no potential keys)
        k3 = mr.split(keys, 3)
        keys = k3[:, 0].contiguous()
        T = float(scenario.temperature)
        pp, lp, gp, gl = evaluate(scenario, initial_state.value, T, k3[:, 1].contiguous())
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
        """Advance the batched state in place by n_iter iterations (one kernel launch for a scenario
        with a fused kernel; per leapfrog step two kick launches around the scenario's potentials
        otherwise)."""
        if not hasattr(scenario, "kernel_args"):
            return self._iterate_user(scenario, state, extra, n_iter, traces or {})
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

    def _iterate_user(self, scenario, state: cdict, extra: cdict, n_iter: int, traces: dict):
        """The fused kernel's iteration for a scenario without one: mocat_hmc_begin (always and the
        proposal's keys), per leapfrog step kick(drift) -> the scenario's potentials and gradients at
        the new position (with that step's key, as mocat's leapfrog, mocat/src/utils.py:117-146) ->
        kick, then mocat_hmc_correct. Same key tree and update arithmetic as the fused kernel."""
        C, d = state.value.shape
        p = extra.parameters
        L = int(p.leapfrog_steps)
        T = float(scenario.temperature)
        mode, target, rm_scale, rm_exp, cut = self._correction_args()
        dev = state.value.device
        mom0 = torch.empty_like(state.momenta)
        kin0 = torch.empty(C, dtype=torch.float32, device=dev)
        step_keys = torch.empty((C, max(L, 1), 2), dtype=torch.uint32, device=dev)
        st = _lib.stream_handle()
        for it in range(int(n_iter)):
            _lib.call("mocat_hmc_begin", C, d, L, float(p.friction), _lib.ptr(state.stepsize), _lib.ptr(state.momenta),
                      _lib.ptr(mom0), _lib.ptr(kin0), _lib.ptr(extra.random_key), _lib.ptr(step_keys), st)
            x1 = state.value.clone()
            g1 = state.grad_potential.clone()
            pot1, prior1, lik1 = state.potential.clone(), state.prior_potential.clone(), \
                state.likelihood_potential.clone()
            for s_ in range(L):
                _lib.call("mocat_hmc_kick", C, d, _lib.ptr(x1), _lib.ptr(state.momenta), _lib.ptr(g1),
                          _lib.ptr(state.stepsize), 1, st)
                prior1, lik1, gp, gl = evaluate(scenario, x1, T, step_keys[:, s_].contiguous())
                pot1 = (prior1 + T * lik1).contiguous()
                g1 = (gp + T * gl).contiguous()
                _lib.call("mocat_hmc_kick", C, d, None, _lib.ptr(state.momenta), _lib.ptr(g1),
                          _lib.ptr(state.stepsize), 0, st)
            extra.iter += 1
            _lib.call("mocat_hmc_correct", C, d, int(extra.iter), mode, target, rm_scale, rm_exp, cut,
                      _lib.ptr(state.value), _lib.ptr(state.grad_potential), _lib.ptr(state.potential),
                      _lib.ptr(state.prior_potential), _lib.ptr(state.likelihood_potential), _lib.ptr(x1), _lib.ptr(g1),
                      _lib.ptr(pot1), _lib.ptr(prior1), _lib.ptr(lik1),
                      _lib.ptr(state.momenta), _lib.ptr(mom0), _lib.ptr(kin0), _lib.ptr(state.stepsize),
                      _lib.ptr(state.alpha), _lib.ptr(extra.random_key), st)
            for k, v in traces.items():
                v[it].copy_(getattr(state, k))
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
