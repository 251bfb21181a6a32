"""Backward simulation for state-space models (mocat/src/ssm/backward.py): forward filtering then
full backward resampling of whole trajectories on the GPU.

Full backward resampling is one launch of the FFBSi kernel per step (one wave per trajectory,
Gumbel-max over the N filter particles with the same key tree as mocat: t_keys = split(key, T),
then split(t_keys[t], M) per trajectory), N x M transition potentials per step. Rejection sampling
(maximum_rejections > 0) runs mocat's accept/reject rounds on the GPU, two launches per round.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .. import _lib
from .. import random as mr
from ..core import cdict
from .filtering import run_particle_filter_for_marginals

_P = _lib.c_ptr
_lib.register("mocat_ffbsi_step", [_lib.c_i32, _lib.c_i32, _P, _P, _P, _lib.c_i64, _P, _lib.c_i64, _P, _P, _P, _P])
_lib.register("mocat_resample_categorical", [_P, _lib.c_i64, _lib.c_i64, _P, _P, _P])
_lib.register("mocat_gather_rows", [_P, _P, _lib.c_i64, _lib.c_i64, _P, _P])
_lib.register("mocat_ffbsi_rejection_workspace_bytes", [_lib.c_i64, _P])
_lib.register("mocat_ffbsi_rejection_step", [_lib.c_i32, _lib.c_i32, _P, _P, _P, _lib.c_i64, _P, _lib.c_i64, _P,
                                             _lib.c_i32, _lib.c_f32, _lib.c_f32, _P, _lib.c_i64, _P, _P, _P, _P])


def backward_simulation(ssm_scenario, marginal_particles: cdict, random_key, n_samps: int = None,
                        maximum_rejections: int = 0, transition_dens_bound_parameter: float = 0.,
                        bound_inflation: float = 1.01) -> cdict:
    """mocat/src/ssm/backward.py:246-270: full backward resampling (maximum_rejections = 0,
    backward_simulation_full :210-243) or rejection sampling against an adaptive density bound
    (backward_simulation_rejection :197-207). Returns a copy of marginal_particles without
    log_weight, with value (T, n_samps, d) the backward trajectories, num_transition_evals (T,)
    (mocat's layout: a leading 0, then the steps from t = 0) and ancestors (T, n_samps): the filter
    particle each trajectory takes at each time."""
    model = getattr(ssm_scenario, "kernel_model", None)
    if model is None or not hasattr(ssm_scenario, "transition_params"):
        raise NotImplementedError("backward simulation is implemented for the GPU state-space models")
    _lib.require_gpu()
    vals = torch.as_tensor(marginal_particles.value, dtype=torch.float32, device="cuda").contiguous()
    lw = torch.as_tensor(marginal_particles.log_weight, dtype=torch.float32, device="cuda").contiguous()
    T, n, d = vals.shape
    m = n if n_samps is None else int(n_samps)
    if n * m >= 0xFFFFFFFF:
        raise ValueError("the final-time draw uses an (n_samps, n) Gumbel matrix: n * n_samps must be < 2^32 - 1")
    if model == 0 and d > 16:
        raise NotImplementedError("backward simulation runs the linear-Gaussian model up to d = 16")
    if model == 1 and hasattr(marginal_particles, "t") and T > 1:
        tt = torch.as_tensor(marginal_particles.t).double().cpu().numpy()
        if not np.allclose(np.diff(tt), 1.0):
            raise NotImplementedError("the stochastic-volatility transition assumes unit time steps")
    params = torch.as_tensor(ssm_scenario.transition_params(), device="cuda").contiguous()
    keys = mr.split(random_key, T).contiguous()
    out = torch.empty((T, m, d), dtype=torch.float32, device="cuda")
    anc = torch.empty((T, m), dtype=torch.int32, device="cuda")
    st = _lib.stream_handle()
    # final time: random.categorical(t_keys[-1], log_weight[-1], shape=(n_samps,)) (backward.py:222-224)
    _lib.call("mocat_resample_categorical", _lib.ptr(lw[T - 1]), n, m, _lib.ptr(keys[T - 1]), _lib.ptr(anc[T - 1]),
              st)
    _lib.call("mocat_gather_rows", _lib.ptr(vals[T - 1]), _lib.ptr(anc[T - 1]), m, d, _lib.ptr(out[T - 1]), st)
    if maximum_rejections > 0:
        nb = _lib.ctypes.c_int64(0)
        _lib.call("mocat_ffbsi_rejection_workspace_bytes", m, _lib.ctypes.byref(nb))
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        evals = torch.zeros(T, dtype=torch.int64, device="cuda")
        for ind in range(T - 2, -1, -1):
            _lib.call("mocat_ffbsi_rejection_step", model, d, _lib.ptr(params), _lib.ptr(vals[ind]),
                      _lib.ptr(lw[ind]), n, _lib.ptr(out[ind + 1]), m, _lib.ptr(keys[ind]), int(maximum_rejections),
                      float(transition_dens_bound_parameter), float(bound_inflation), _lib.ptr(ws), int(ws.numel()),
                      _lib.ptr(anc[ind]), _lib.ptr(out[ind]), _lib.ptr(evals[ind + 1]), st)
        num_evals = evals.cpu().to(torch.float64)  # [0, steps t = 0 .. T-2] (backward.py:205)
    else:
        for ind in range(T - 2, -1, -1):
            _lib.call("mocat_ffbsi_step", model, d, _lib.ptr(params), _lib.ptr(vals[ind]), _lib.ptr(lw[ind]), n,
                      _lib.ptr(out[ind + 1]), m, _lib.ptr(keys[ind]), _lib.ptr(anc[ind]), _lib.ptr(out[ind]), st)
        num_evals = torch.cat([torch.zeros(1, dtype=torch.float64),
                               torch.full((T - 1,), float(n) * m, dtype=torch.float64)])
    res = cdict(**{k: v for k, v in marginal_particles.__dict__.items() if k != "log_weight"})
    res.value = out
    res.ancestors = anc
    res.num_transition_evals = num_evals
    return res


def forward_filtering_backward_simulation(ssm_scenario, particle_filter, y, t, n_samps: int, random_key,
                                          n_pf: int = None, ess_threshold: float = 0.5, maximum_rejections: int = 0,
                                          transition_dens_bound_parameter: float = 0., bound_inflation: float = 1.01,
                                          resampling: str = "categorical") -> cdict:
    """mocat/src/ssm/backward.py:273-315: pf_key, bsi_key = split(key); a bootstrap filter of n_pf
    (default n_samps) particles, then backward simulation; time, pf_time and bsi_time in seconds."""
    keys = mr.split(random_key, 2)
    if n_pf is None:
        n_pf = n_samps
    torch.cuda.synchronize()
    pf_start = time.perf_counter()
    pf_samps = run_particle_filter_for_marginals(ssm_scenario, particle_filter, y, t, keys[0], n=n_pf,
                                                 ess_threshold=ess_threshold, resampling=resampling)
    torch.cuda.synchronize()
    pf_end = time.perf_counter()
    bsi = backward_simulation(ssm_scenario, pf_samps, keys[1], n_samps, maximum_rejections,
                              transition_dens_bound_parameter, bound_inflation)
    torch.cuda.synchronize()
    bsi_end = time.perf_counter()
    bsi.time = bsi_end - pf_start
    bsi.pf_time = pf_end - pf_start
    bsi.bsi_time = bsi_end - pf_end
    return bsi
