"""Transport samplers (mocat/src/transport/{sampler,smc}.py) on the HIP kernels.

Tempered SMC (MetropolisedSMCSampler / RMMetropolisedSMCSampler, mocat/src/transport/smc.py:228-428)
per temperature iteration, as mocat orders it (smc.py:73-99, 193-225):
  keys = split(key, n + 2); key = keys[-1]
  resample with keys[-2] if ESS <= ess_threshold_resample * n   (gather of every particle field)
  move every particle at the current temperature: Underdamped.startup (re-evaluates potential and
      gradient, key split(keys[i], 3)[0]) then mcmc_steps Metropolis-corrected updates
      (fused iteration kernel; MALA when leapfrog_steps = 1)
  next temperature by regula falsi on the ESS (FP64, one workgroup), reweight, ESS, log-normaliser,
      Robbins-Monro stepsize update from the weighted mean acceptance (RM variant)
The iteration stops when the temperature reaches max_temperature.

Difference from mocat: the moved likelihood potential is taken from the kernel, which tracks it
separately, instead of (potential - prior) / temperature (smc.py:363-364), which loses precision at
small temperatures.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _lib
from . import random as mr
from .core import Scenario, cdict, is_implemented
from .mcmc import Metropolis, check_correction, evaluate
from .sample import Sampler

_P = _lib.c_ptr
_lib.register("mocat_gather_rows", [_P, _P, _lib.c_i64, _lib.c_i64, _P, _P])
_lib.register("mocat_smc_temper", [_P, _P, _lib.c_i64, _lib.c_f64, _lib.c_f64, _lib.c_f64, _lib.c_f64, _lib.c_f64,
                                   _lib.c_i32, _lib.c_i32, _lib.c_f64, _P, _lib.c_i32, _lib.c_f64, _lib.c_f64,
                                   _lib.c_f64, _P, _P, _P])


def gather_rows(t: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """t[idx] along the first axis (4-byte element types)."""
    t = t.contiguous()
    if t.element_size() != 4:
        raise TypeError("gather_rows handles 4-byte element types")
    out = torch.empty((idx.shape[0], *t.shape[1:]), dtype=t.dtype, device=t.device)
    width = int(t[0].numel()) if t.dim() > 1 else 1
    _lib.call("mocat_gather_rows", _lib.ptr(t), _lib.ptr(idx), int(idx.shape[0]), width, _lib.ptr(out),
              _lib.stream_handle())
    return out


class TransportSampler(Sampler):
    def startup_particles(self, scenario, n, initial_state, key):
        """TransportSampler.startup (mocat/src/transport/sampler.py:18-34)."""
        if initial_state is None:
            s = mr.split(key, 2)
            key, sub = s[0].contiguous(), s[1].contiguous()
            if is_implemented(scenario.prior_sample):
                vals = scenario.prior_sample(mr.split(sub, n))
            else:
                vals = mr.normal(sub, (n, scenario.dim))
            initial_state = cdict(value=vals.contiguous())
        return initial_state, key


class SMCSampler(TransportSampler):
    name = "SMC Sampler"


class TemperedSMCSampler(SMCSampler):
    name = "Tempered SMC Sampler"


class MetropolisedSMCSampler(TemperedSMCSampler):
    name = "Metropolised SMC Sampler"
    rm = False

    def __init__(self, mcmc_sampler, mcmc_correction="sampler_default", mcmc_steps: int = 1,
                 max_iter: int = int(1e4), temperature_schedule=None, max_temperature: float = 1.0,
                 ess_threshold_retain: float = 0.9, ess_threshold_resample: float = 0.5,
                 bisection_tol: float = 1e-5, max_bisection_iter: int = 1000, resampling: str = "categorical",
                 **kwargs):
        super().__init__(**kwargs)
        if isinstance(mcmc_sampler, type):
            mcmc_sampler = mcmc_sampler()
        self.mcmc_sampler = mcmc_sampler
        if mcmc_correction != "sampler_default":
            self.mcmc_sampler.correction = mcmc_correction
        if temperature_schedule is not None:
            ts = np.asarray(temperature_schedule, np.float64)
            if ts[0] == 0.0:
                ts = ts[1:]
            temperature_schedule = ts
            max_temperature = float(ts[-1])
            max_iter = len(ts)
        self.temperature_schedule = temperature_schedule
        self.max_temperature = max_temperature
        self.max_iter = int(max_iter)
        self.resampling = resampling
        self.parameters.mcmc_steps = int(mcmc_steps)
        self.parameters.ess_threshold_retain = ess_threshold_retain
        self.parameters.ess_threshold_resample = ess_threshold_resample
        self.parameters.bisection_tol = bisection_tol
        self.parameters.max_bisection_iter = int(max_bisection_iter)

    # -- device helpers -------------------------------------------------------------------------
    def _temper(self, lw, lik, beta, ess, it, log_h, alpha_trace=None):
        p = self.parameters
        out = torch.empty(8, dtype=torch.float64, device=lw.device)
        lw_out = torch.empty_like(lw)
        adaptive = self.temperature_schedule is None
        nb = 0.0 if adaptive else float(self.temperature_schedule[min(it, len(self.temperature_schedule) - 1)])
        k_alpha = 0 if alpha_trace is None else int(alpha_trace.shape[0])
        rm = float(p.rm_stepsize) if (self.rm and alpha_trace is not None) else 0.0
        target = float(self.mcmc_sampler.tuning.target or 0.0)
        _lib.call("mocat_smc_temper", _lib.ptr(lw), _lib.ptr(lik), int(lw.shape[0]), float(beta), float(ess),
                  float(p.ess_threshold_retain), float(self.max_temperature), float(p.bisection_tol),
                  int(p.max_bisection_iter), int(adaptive), nb, _lib.ptr(alpha_trace), k_alpha, float(log_h), rm,
                  target, _lib.ptr(lw_out), _lib.ptr(out), _lib.stream_handle())
        return lw_out, out.cpu().numpy()

    def _resample(self, state, lw, key):
        n = lw.shape[0]
        if self.resampling == "categorical":
            from .ssm.filtering import categorical
            idx = categorical(key, lw, n)
        else:
            from .ssm.filtering import resample_indices
            idx = resample_indices(lw, key, self.resampling, log_weights=True)
        for k in ("value", "grad_potential", "momenta", "potential", "prior_potential", "likelihood_potential",
                  "alpha"):
            setattr(state, k, gather_rows(getattr(state, k), idx))
        return state, idx

    def _user_moves(self, scenario, state, stepsize, keys, beta, steps, alpha_trace):
        """forward_proposal's corrected Underdamped updates at temperature beta for a scenario without a
        fused kernel: the MCMC sampler's iteration in pieces (its potentials between launches), the
        same key tree and arithmetic as the fused moves (Metropolis, fixed stepsize, extra.iter from 0)."""
        ms = self.mcmc_sampler
        p = self.parameters
        saved, saved_T = ms.correction, scenario.temperature
        ms.correction = Metropolis()
        scenario.temperature = beta
        try:
            st = cdict(value=state.value, momenta=state.momenta, grad_potential=state.grad_potential,
                       potential=state.potential, prior_potential=state.prior_potential,
                       likelihood_potential=state.likelihood_potential, stepsize=stepsize, alpha=state.alpha)
            ex = cdict(random_key=keys, iter=0,
                       parameters=cdict(leapfrog_steps=int(p.leapfrog_steps), friction=float(p.friction)))
            ms._iterate_user(scenario, st, ex, steps, {"alpha": alpha_trace})
        finally:
            ms.correction, scenario.temperature = saved, saved_T
        for k in ("potential", "prior_potential", "likelihood_potential"):
            setattr(state, k, getattr(st, k))

    def run_fused(self, scenario, n, initial_state, initial_extra, n_chains=None,
                  stack=("value", "log_weight"), **kwargs):
        fused = hasattr(scenario, "kernel_args")
        ms = self.mcmc_sampler
        ms.correction = check_correction(ms.correction)
        if not isinstance(ms.correction, Metropolis):
            raise NotImplementedError("SMC moves use a Metropolis-corrected Underdamped kernel")
        p = self.parameters
        for k, v in ms.parameters.__dict__.items():
            if not hasattr(p, k) or getattr(p, k) is None:
                setattr(p, k, v)
        if p.stepsize is None:
            raise ValueError("the MCMC sampler needs a stepsize")
        summary = self.summary(scenario, initial_state, initial_extra)
        key = initial_extra.random_key
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # ---- startup (TransportSampler / TemperedSMCSampler / MetropolisedSMCSampler.startup)
        state, key = self.startup_particles(scenario, n, initial_state, key)
        key = mr.split(key, 2 * n + 1)[-1].contiguous()
        pp, lp, gp, gl = evaluate(scenario, state.value, 0.0)
        state.prior_potential, state.likelihood_potential = pp, lp
        lw0 = torch.zeros(n, dtype=torch.float32, device=pp.device)
        ess = float(n)
        lw, out = self._temper(lw0, lp, 0.0, ess, 0, math.log(p.stepsize))
        beta = float(out[0])
        lnc = float(out[2]) - math.log(n)      # logsumexp(lw, b=1/n)
        ess = float(out[1])
        key = mr.split(key, 3)[0].contiguous()  # vmapped Underdamped.startup
        pp, lp, gp, gl = evaluate(scenario, state.value, beta)
        state.prior_potential, state.likelihood_potential = pp, lp
        state.potential = pp + beta * lp
        state.grad_potential = (gp + beta * gl).contiguous()
        state.momenta = torch.zeros_like(state.value)
        state.alpha = torch.ones(n, dtype=torch.float32, device=pp.device)
        h = float(p.stepsize)
        hist = {k: [] for k in ("temperature", "ess", "log_norm_constant", "stepsize", "alpha_mean",
                                "bisection_iters")}
        stacks = {k: [] for k in stack}

        def record(alpha_mean, iters):
            hist["temperature"].append(beta)
            hist["ess"].append(ess)
            hist["log_norm_constant"].append(lnc)
            hist["stepsize"].append(h)
            hist["alpha_mean"].append(alpha_mean)
            hist["bisection_iters"].append(iters)
            for k in stack:
                stacks[k].append((lw if k == "log_weight" else getattr(state, k)).clone())

        record(1.0, 0)
        it = 0
        if fused:
            X, y, packed, nrows, d, prec, pprec = scenario.kernel_args()
        steps = int(p.mcmc_steps)
        alpha_trace = torch.empty((steps, n), dtype=torch.float32, device=pp.device)
        while beta < self.max_temperature and it < self.max_iter:
            it += 1
            keys_all = mr.split(key, n + 2)
            key = keys_all[-1].contiguous()
            if ess <= p.ess_threshold_resample * n:
                state, _ = self._resample(state, lw, keys_all[-2].contiguous())
                lw = torch.zeros(n, dtype=torch.float32, device=lw.device)
                ess = float(n)
            # forward_proposal: Underdamped.startup then mcmc_steps corrected updates at `beta`
            pkeys = mr.split(keys_all[:n].contiguous(), 3)[:, 0].contiguous()
            pp, lp, gp, gl = evaluate(scenario, state.value, beta)
            state.potential = pp + beta * lp
            state.grad_potential = (gp + beta * gl).contiguous()
            state.prior_potential, state.likelihood_potential = pp, lp
            stepsize = torch.full((n,), h, dtype=torch.float32, device=lw.device)
            if not fused:
                self._user_moves(scenario, state, stepsize, pkeys, beta, steps, alpha_trace)
            else:
                _lib.call("mocat_hmc_lr_iterate", X, y, packed, nrows, d, prec, pprec, float(beta), n,
                          _lib.ptr(state.value), _lib.ptr(state.momenta), _lib.ptr(state.grad_potential),
                          _lib.ptr(state.potential), _lib.ptr(state.prior_potential),
                          _lib.ptr(state.likelihood_potential), _lib.ptr(stepsize), _lib.ptr(state.alpha),
                          _lib.ptr(pkeys), 0, steps, int(p.leapfrog_steps), float(p.friction), 1, 0.0, 1.0, 2 / 3,
                          math.inf, _P(0), _P(0), _lib.ptr(alpha_trace), _P(0), _P(0), _lib.stream_handle())
            # adapt: next temperature, reweight, ESS, log-normaliser, RM stepsize
            lw, out = self._temper(lw, state.likelihood_potential, beta, ess, it, math.log(h), alpha_trace)
            beta_new = float(out[0])
            lnc += float(out[2]) - float(out[3])
            ess = float(out[1])
            if self.rm:
                h = math.exp(float(out[4]))
            beta = beta_new
            state.potential = state.prior_potential + beta * state.likelihood_potential
            state.alpha = alpha_trace.mean(0) if steps > 0 else state.alpha
            record(float(out[5]), int(out[6]))
            if bool(torch.isnan(state.value).float().mean() > 0.1):
                break
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        chain = cdict()
        for k in stack:
            chain.__dict__[k] = torch.stack(stacks[k])
        dev = lw.device
        for k, v in hist.items():
            chain.__dict__[k] = torch.tensor(v, dtype=torch.float64, device=dev)
        chain.final_state = state
        chain.final_log_weight = lw
        scenario.temperature = beta
        return chain, elapsed, summary


class RMMetropolisedSMCSampler(MetropolisedSMCSampler):
    rm = True

    def __init__(self, mcmc_sampler, mcmc_correction=Metropolis, mcmc_steps: int = 1, max_iter: int = int(1e4),
                 temperature_schedule=None, max_temperature: float = 1.0, ess_threshold_retain: float = 0.9,
                 ess_threshold_resample: float = 0.5, bisection_tol: float = 1e-5, max_bisection_iter: int = 1000,
                 rm_stepsize: float = 1.0, resampling: str = "categorical", **kwargs):
        super().__init__(mcmc_sampler, mcmc_correction, mcmc_steps, max_iter, temperature_schedule, max_temperature,
                         ess_threshold_retain, ess_threshold_resample, bisection_tol, max_bisection_iter,
                         resampling, **kwargs)
        self.parameters.rm_stepsize = rm_stepsize


_lib.register("mocat_svgd_gmm_run_workspace_bytes", [_lib.c_i64, _lib.c_i64, _lib.c_i32, _P])
_lib.register("mocat_svgd_gmm_run", [_P, _lib.c_i64, _lib.c_i64, _lib.c_i32, _P, _P, _P, _lib.c_f32, _lib.c_f32,
                                     _lib.c_f32, _lib.c_i32, _P, _lib.c_i32, _lib.c_i32, _P, _lib.c_i64, _P, _P,
                                     _P, _P, _P, _P, _P])
_lib.register("mocat_svgd_step", [_P, _P, _lib.c_i64, _lib.c_i64, _lib.c_f32, _P, _lib.c_i32, _P, _lib.c_f32,
                                  _lib.c_f32, _P, _P, _P, _P, _P])


class SVGD(TransportSampler):
    """Stein variational gradient descent (mocat/src/transport/svgd.py:42-146), full ensemble batch,
    Gaussian RBF kernel, adagrad(step_size, momentum=0.9) as mocat's default optimiser.

    bandwidth_update: "mean" (mocat's default when no kernel is given, svgd.py:59-61), "median"
    (kernels.median_bandwidth_update), or None (fixed kernel bandwidth).
    j_splits: workgroups sharing the pairwise sum of one 128-particle tile (partial sums combined in
    the update kernel); None picks enough to give every compute unit one workgroup.
    """
    name = "SVGD"

    def __init__(self, stepsize, max_iter: int = 1000, kernel=None, kernel_params: cdict = None,
                 ensemble_batchsize: int = None, optimiser: str = "adagrad", momentum: float = 0.9,
                 bandwidth_update: str = "default", j_splits: int = None, **kwargs):
        super().__init__(max_iter=max_iter, **kwargs)
        from .kernels import Gaussian
        if bandwidth_update == "default":
            bandwidth_update = "mean" if kernel is None else None
        if kernel is None:
            kernel = Gaussian()
        if not isinstance(kernel, Gaussian):
            raise NotImplementedError("the fused SVGD kernel implements the Gaussian RBF kernel")
        if optimiser != "adagrad":
            raise NotImplementedError("the fused SVGD update implements adagrad (mocat's default optimiser)")
        if ensemble_batchsize is not None:
            raise NotImplementedError("the fused SVGD kernel uses the full ensemble (ensemble_batchsize=None)")
        self.kernel = kernel
        self.parameters.stepsize = stepsize
        self.parameters.kernel_params = kernel_params if kernel_params is not None else kernel.parameters
        self.parameters.momentum = momentum
        self.bandwidth_update = bandwidth_update
        self.j_splits = None if j_splits is None else int(j_splits)

    def _js(self, n: int) -> int:
        if self.j_splits is not None:
            return self.j_splits
        tiles, blocks = -(-n // 128), -(-n // 32)
        return max(1, min(32, blocks, -(-256 // tiles)))

    def _adapt(self, x):
        """The kernel bandwidth for the next update: a device scalar when adapted (no host sync), else
        the fixed kernel parameter."""
        from .kernels import mean_bandwidth_update, median_bandwidth_update
        if self.bandwidth_update == "median":
            return median_bandwidth_update(x)
        if self.bandwidth_update == "mean":
            return mean_bandwidth_update(x)
        return float(self.parameters.kernel_params.bandwidth)

    def step(self, state: cdict, extra: cdict, scenario, workspace=None):
        """One SVGD.update (svgd.py:122-146) in place on the device state (value, grad_potential,
        potential, adagrad g_sq / m); returns phi."""
        x, g = state.value, state.grad_potential
        n, d = x.shape
        extra.iter += 1
        ks = mr.split(extra.random_key, n + 2)
        extra.random_key = ks[-2].contiguous()
        step = self.parameters.stepsize
        step = float(step(extra.iter)) if callable(step) else float(step)
        if workspace is None:
            from .kernels import _workspace
            workspace = _workspace(n, d, self._js(n), x.device)
        x_in = x.clone()
        phi = torch.empty_like(x)
        bw = extra.bandwidth
        if isinstance(bw, torch.Tensor):
            bw_dev, bw_host = bw.to(device=x.device, dtype=torch.float32).reshape(1).contiguous(), 1.0
        else:
            bw_dev, bw_host = None, float(bw)
        _lib.call("mocat_svgd_step", _lib.ptr(x_in), _lib.ptr(g), n, d, bw_host, _lib.ptr(bw_dev), self._js(n),
                  _lib.ptr(workspace), step, float(self.parameters.momentum), _lib.ptr(x), _lib.ptr(extra.g_sq),
                  _lib.ptr(extra.m), _lib.ptr(phi), _lib.stream_handle())
        T = float(scenario.temperature)
        pp, lp, gp, gl = evaluate(scenario, x, T)
        state.potential = pp + T * lp
        state.grad_potential = (gp + T * gl).contiguous()
        extra.bandwidth = self._adapt(x)
        return phi

    def run_fused(self, scenario, n, initial_state, initial_extra, n_chains=None, stack=(), **kwargs):
        summary = self.summary(scenario, initial_state, initial_extra)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        state, key = self.startup_particles(scenario, n, initial_state, initial_extra.random_key)
        x = state.value.to(torch.float32).contiguous()
        key = mr.split(key, n + 1)[-1].contiguous()
        T = float(scenario.temperature)
        pp, lp, gp, gl = evaluate(scenario, x, T)
        state = cdict(value=x, potential=pp + T * lp, grad_potential=(gp + T * gl).contiguous())
        extra = cdict(random_key=key, iter=0, g_sq=torch.zeros_like(x), m=torch.zeros_like(x))
        extra.bandwidth = float(self.parameters.kernel_params.bandwidth)
        extra.bandwidth = self._adapt(x)
        from .kernels import _workspace
        from .scenarios import GaussianMixture
        if (isinstance(scenario, GaussianMixture) and not scenario.full and not stack
                and not callable(self.parameters.stepsize) and n >= 3):
            return self._run_gmm_fused(scenario, state, extra, T, summary, t0)
        ws = _workspace(n, x.shape[1], self._js(n), x.device)
        stacks = {k: [getattr(state, k).clone()] for k in stack}
        bws = [extra.bandwidth]
        for _ in range(self.max_iter):
            self.step(state, extra, scenario, ws)
            bws.append(extra.bandwidth)
            for k in stack:
                stacks[k].append(getattr(state, k).clone())
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        chain = cdict()
        for k in stack:
            chain.__dict__[k] = torch.stack(stacks[k])
        chain.bandwidth = torch.stack([b.reshape(()).to(torch.float32) if isinstance(b, torch.Tensor)
                                       else torch.tensor(b, dtype=torch.float32, device=x.device) for b in bws])
        chain.final_state = state
        chain.final_extra = extra
        return chain, elapsed, summary

    def _run_gmm_fused(self, scenario, state, extra, T, summary, t0):
        """All iterations in one library call (Gaussian-mixture targets): the same updates, in the same
        order, as repeated step() calls, with no host round trip between iterations."""
        x = state.value
        n, d = x.shape
        js = self._js(n)
        mode = {"median": 2, "mean": 1, None: 0}[self.bandwidth_update]
        bw = extra.bandwidth
        bw = (bw.to(device=x.device, dtype=torch.float32).reshape(1).clone() if isinstance(bw, torch.Tensor)
              else torch.full((1,), float(bw), dtype=torch.float32, device=x.device))
        bw0 = bw.clone()
        nb = _lib.ctypes.c_int64(0)
        _lib.call("mocat_svgd_gmm_run_workspace_bytes", n, d, js, _lib.ctypes.byref(nb))
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device=x.device)
        trace = torch.empty(self.max_iter, dtype=torch.float32, device=x.device)
        pot = state.potential.to(torch.float32).contiguous()
        grad = state.grad_potential.to(torch.float32).contiguous()
        key = extra.random_key.clone().contiguous()
        _lib.call("mocat_svgd_gmm_run", _lib.ptr(x), n, d, scenario.n_components, _lib.ptr(scenario.means),
                  _lib.ptr(scenario.lam), _lib.ptr(scenario.logc), float(T), float(self.parameters.stepsize),
                  float(self.parameters.momentum), mode, _lib.ptr(bw), int(self.max_iter), js, _lib.ptr(ws),
                  int(ws.numel()), _lib.ptr(extra.g_sq), _lib.ptr(extra.m), _lib.ptr(pot), _lib.ptr(grad),
                  _lib.ptr(key), _lib.ptr(trace), _lib.stream_handle())
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        state.potential, state.grad_potential = pot, grad
        extra.iter += self.max_iter
        extra.random_key = key
        extra.bandwidth = bw[0]
        chain = cdict()
        chain.bandwidth = torch.cat([bw0, trace])
        chain.final_state = state
        chain.final_extra = extra
        return chain, elapsed, summary
