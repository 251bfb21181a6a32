"""Particle filtering (mocat/src/ssm/filtering.py) on the persistent HIP filter kernel.

run_particle_filter_for_marginals keeps mocat's signature and key tree
(mocat/src/ssm/filtering.py:255-324) and output cdict(value (T, n, d), log_weight (T, n), y, t,
ess (T,)), and adds: `resampling` ("categorical" = mocat's Gumbel-max, the default; "systematic";
"multinomial"), `log_likelihood` (FP64 estimate of log p(y_0:T)) and `ancestors` (T, n).

run_particle_filters runs many independent filters of uneven particle counts and lengths in one
launch (one workgroup per filter, largest n*T first) and returns their FP64 log-likelihoods.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .. import random as mr
from ..core import cdict

LOG2PI = float(np.log(2 * np.pi))

RESAMPLING = {"systematic": 0, "multinomial": 1, "categorical": 2}

_P = _lib.c_ptr
_lib.register("mocat_pf_param_count", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _P])
_lib.register("mocat_pf_run", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _P, _P, _lib.c_i32, _P, _P, _P, _P, _P, _P, _P, _P,
                               _P, _lib.c_i32, _lib.c_f32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("mocat_pf_large_workspace_bytes", [_P, _lib.c_i32, _P])
_lib.register("mocat_pf_run_large", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _P, _lib.c_i32, _P, _P, _P, _P, _P, _P, _P,
                                     _P, _lib.c_i32, _lib.c_f32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _lib.c_i64,
                                     _lib.c_i32, _P])
_lib.register("mocat_pf_hook", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _lib.c_i32, _P, _P, _P, _P, _lib.c_f32, _lib.c_i64,
                                _P, _P, _P])
_lib.register("mocat_resample", [_P, _lib.c_i64, _lib.c_i32, _lib.c_i32, _P, _P, _P, _P])
_lib.register("mocat_pf_compiled_shape", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _P])
_lib.register("mocat_gather_rows", [_P, _P, _lib.c_i64, _lib.c_i64, _P, _P])
_lib.register("mocat_enkf_hook", [_lib.c_i32, _lib.c_i32, _lib.c_i32, _P, _P, _P, _lib.c_i64, _P, _P, _P, _P])
_lib.register("mocat_resample_noalloc", [_P, _lib.c_i64, _lib.c_i32, _lib.c_i32, _P, _P, _P])
_lib.register("mocat_resample_categorical", [_P, _lib.c_i64, _lib.c_i64, _P, _P, _P])
_lib.register("mocat_enkf_param_count", [_lib.c_i32, _lib.c_i32, _P])
_lib.register("mocat_enkf_workspace_bytes", [_lib.c_i64, _lib.c_i32, _P])
_lib.register("mocat_enkf_run", [_lib.c_i32, _lib.c_i32, _P, _P, _lib.c_i64, _lib.c_i64, _P, _P, _P, _lib.c_i64, _P, _P,
                                 _P])

FILTER_DTYPE = np.dtype([("n", "<i8"), ("T", "<i8"), ("y_offset", "<i8"), ("state_offset", "<i8"),
                         ("out_offset", "<i8"), ("t_offset", "<i8"), ("key", "<u4", (2,)),
                         ("param_offset", "<i4"), ("flags", "<i4")])
PF_LARGE = 1  # flags bit: the filter runs on the chunked many-workgroup path (mocat_pf_run_large)
PF_RESUME = 2  # flags bit: start from given particles and log-weights (initial_sample)
PF_STEP_KEY = 4  # flags bit: the key of step 1 is the filter key itself (propagate_particle_filter)
PF_INIT_KEY = 8  # flags bit: initial particles from split(key, n) (initiate_particles)
assert FILTER_DTYPE.itemsize == 64


HOOK_ERROR = ("{what}: the particle-filter hooks run on compiled models only (TimeHomogenousLinearGaussian and "
              "StochasticVolatility under BootstrapFilter; TimeHomogenousLinearGaussian under "
              "OptimalNonLinearGaussianParticleFilter). A StateSpaceModel or ParticleFilter written in Python has "
              "no kernel on the fused path.")


def _hook_target(particle_filter, ssm_scenario):
    """(kernel model id, parameter vector) of the batched hooks, or NotImplementedError (HOOK_ERROR)."""
    kind = BootstrapFilter if particle_filter is None else None
    if kind is None:
        try:
            kind = _filter_kind(particle_filter)
        except NotImplementedError:
            raise NotImplementedError(HOOK_ERROR.format(what=type(particle_filter).__name__)) from None
    if kind is EnsembleKalmanFilter:
        raise NotImplementedError("EnsembleKalmanFilter: its analysis step needs the whole ensemble's covariance; "
                                  "run it with run_particle_filter_for_marginals")
    pf = particle_filter if isinstance(particle_filter, ParticleFilter) else kind()
    # a subclass that overrides a per-particle rule would not be honoured by the compiled hook
    for name in ("proposal_sample", "intermediate_log_weight", "initial_sample", "initial_log_weight"):
        if getattr(type(pf), name, None) is not getattr(kind, name, None):
            raise NotImplementedError(HOOK_ERROR.format(what=f"{type(pf).__name__}.{name}"))
    if kind is OptimalNonLinearGaussianParticleFilter:
        try:
            return 2, pf.kernel_params(ssm_scenario)
        except NotImplementedError:
            raise NotImplementedError(HOOK_ERROR.format(what=type(ssm_scenario).__name__)) from None
    km = getattr(ssm_scenario, "kernel_model", None)
    if km not in (0, 1) or not hasattr(ssm_scenario, "kernel_params"):
        raise NotImplementedError(HOOK_ERROR.format(what=type(ssm_scenario).__name__))
    return km, ssm_scenario.kernel_params()


def _step_scale(ssm_scenario, t_previous, t_new) -> float:
    """The transition's noise scale for a step from t_previous to t_new: sqrt(t_new - t_previous) for
    the stochastic-volatility model (mocat/src/ssm/scenarios/stochastic_volatility.py:86-93), 1 for the
    time-homogeneous linear-Gaussian models and when either time is not given."""
    if getattr(ssm_scenario, "kernel_model", None) != 1 or t_previous is None or t_new is None:
        return 1.0
    dt = float(t_new) - float(t_previous)
    if not dt >= 0.0:
        raise ValueError("observation times must not decrease")
    return float(np.sqrt(np.float32(dt)))


def _scale_rows(ssm_scenario, t, T):
    """Per-row noise scales of one filter (row t: the step into observation t; row 0 unused), or None
    for unit steps or a time-homogeneous model."""
    if getattr(ssm_scenario, "kernel_model", None) != 1 or t is None:
        return None
    tn = np.asarray(t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else t, np.float64).reshape(-1)
    if len(tn) != T:
        raise ValueError(f"expected {T} observation times, got {len(tn)}")
    dt = np.diff(tn)
    if np.any(dt < 0):
        raise ValueError("observation times must not decrease")
    rows = np.ones(T, np.float32)
    rows[1:] = np.sqrt(dt.astype(np.float32))
    return None if np.all(rows == 1.0) else rows


def _hook(model, params, ssm_scenario, phase, x_prev, y, keys, n, step_scale=1.0):
    d, dy = int(ssm_scenario.dim), int(ssm_scenario.dim_obs)
    dev = mr.default_device()
    p = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(dev)
    yt = None if y is None else torch.as_tensor(y, dtype=torch.float32, device=dev).reshape(-1).contiguous()
    if yt is not None and yt.numel() != dy:
        raise ValueError(f"y must hold dim_obs = {dy} values")
    xp = None if x_prev is None else torch.as_tensor(x_prev, dtype=torch.float32, device=dev).reshape(n, d).contiguous()
    kt = None if keys is None else mr.as_key(keys).reshape(-1, 2).contiguous()
    if kt is not None and model != 2 and kt.shape[0] != n:
        raise ValueError(f"expected {n} keys, got {kt.shape[0]}")
    x = torch.empty((n, d), dtype=torch.float32, device=dev) if phase != 2 else None
    lw = torch.empty(n, dtype=torch.float32, device=dev)
    _lib.call("mocat_pf_hook", model, d, dy, phase, _lib.ptr(p), _lib.ptr(xp), _lib.ptr(yt), _lib.ptr(kt),
              float(step_scale), int(n), _lib.ptr(x), _lib.ptr(lw), _lib.stream_handle())
    return x, lw


class ParticleFilter:
    """mocat's ParticleFilter (mocat/src/ssm/filtering.py:20-139). The batched hooks below are the seams
    mocat's scan calls; here they run the compiled models' kernels (one thread per particle, the same
    noise draws as the fused filters), and raise HOOK_ERROR's NotImplementedError for anything else."""
    name = "Particle Filter"

    def __init__(self, name: str = None):
        if name is not None:
            self.name = name

    def __repr__(self):
        return f"mocat_amd.ParticleFilter.{self.__class__.__name__}"

    def startup(self, ssm_scenario):
        pass

    def initial_sample_vectorised(self, ssm_scenario, y, t, random_keys):
        """filtering.py:56-63: one initial sample per key (random_keys (n, 2))."""
        return self.initial_sample_and_weight_vectorised(ssm_scenario, y, t, random_keys)[0]

    def initial_sample_and_weight_vectorised(self, ssm_scenario, y, t, random_keys):
        """filtering.py:65-78: (x (n, d), log_weight (n,)); log-weights are -likelihood(x, y), zeros
        without y."""
        model, params = _hook_target(self, ssm_scenario)
        n = int(mr.as_key(random_keys).reshape(-1, 2).shape[0])
        return _hook(model, params, ssm_scenario, 0, None, y, random_keys, n)

    def proposal_sample_vectorised(self, ssm_scenario, x_previous, t_previous, y_new, t_new, random_keys):
        """filtering.py:113-122."""
        return self.propose_and_intermediate_weight_vectorised(ssm_scenario, x_previous, t_previous, y_new, t_new,
                                                               random_keys)[0]

    def propose_and_intermediate_weight_vectorised(self, ssm_scenario, x_previous, t_previous, y_new, t_new,
                                                   random_keys):
        """filtering.py:124-139: (x_new (n, d), intermediate log-weights (n,)) from x_previous (n, d) and
        one key per particle."""
        model, params = _hook_target(self, ssm_scenario)
        sc = _step_scale(ssm_scenario, t_previous, t_new)
        xp = torch.as_tensor(x_previous)
        n = int(xp.shape[0]) if xp.dim() > 1 else 1
        return _hook(model, params, ssm_scenario, 1, xp, y_new, random_keys, n, sc)


class BootstrapFilter(ParticleFilter):
    """Proposal = transition, weight = likelihood (mocat/src/ssm/filtering.py:142-170)."""
    name = "Bootstrap Filter"

    def initial_sample(self, ssm_scenario, y, t, random_key):
        """filtering.py:40-45 for one key."""
        return ssm_scenario.initial_sample(t, random_key)

    def initial_log_weight(self, ssm_scenario, x, y, t):
        return -ssm_scenario.likelihood_potential(x, y, t)

    def proposal_sample(self, ssm_scenario, x_previous, t_previous, y_new, t_new, random_key):
        """filtering.py:154-161: the transition sample."""
        return ssm_scenario.transition_sample(x_previous, t_previous, t_new, random_key)

    def intermediate_log_weight(self, ssm_scenario, x_previous, t_previous, x_new, y_new, t_new):
        """filtering.py:163-170: -likelihood_potential(x_new, y_new)."""
        return -ssm_scenario.likelihood_potential(x_new, y_new, t_new)


def _sym_chol(M):
    # jnp.linalg.cholesky symmetrises its input, (M + M^T) / 2, before factoring
    return np.linalg.cholesky(0.5 * (M + M.T))


def _kalman_gain(cov_sqrt, H, lik_cov_sqrt, inv_mat=None):
    """mocat/src/utils.py:476-484."""
    if inv_mat is None:
        hc = H @ cov_sqrt
        inv_mat = np.linalg.inv(hc @ hc.T + lik_cov_sqrt @ lik_cov_sqrt.T)
    return cov_sqrt @ cov_sqrt.T @ H.T @ inv_mat


def _nlg_matrices(model):
    """The linear-Gaussian model as mocat's NonLinearGaussian interface reads it: transition_function
    x -> A x (linear_gaussian.py:112-116) and reset_covariance's square roots (utils.py:247-262)."""
    if getattr(model, "kernel_model", None) != 0:
        raise NotImplementedError("the optimal-proposal filter and the EnKF are compiled for the linear transition "
                                  "x -> A x of TimeHomogenousLinearGaussian")
    L0 = np.linalg.cholesky(model.initial_covariance)
    return dict(m0=np.asarray(model.initial_mean, np.float64), L0=L0, S0=np.linalg.inv(L0),
                A=np.asarray(model.transition_matrix, np.float64), Q=np.asarray(model.transition_covariance, np.float64),
                LQ=model.transition_covariance_sqrt, H=np.asarray(model.likelihood_matrix, np.float64),
                LR=model.likelihood_covariance_sqrt, SR=model.likelihood_precision_sqrt)


def _initial_conditioning(M):
    """initial_kalman_gain and initial_conditioned_covariance_sqrt (nonlinear_gaussian.py:161-170,
    304-310), mocat's expressions as written."""
    K0 = _kalman_gain(M["L0"], M["H"], M["LR"])
    hs = M["H"].T @ M["SR"]
    icp = M["S0"] @ M["S0"] + hs @ hs.T
    L = _sym_chol(icp)
    return K0, L, np.linalg.inv(L), icp


class OptimalNonLinearGaussianParticleFilter(ParticleFilter):
    """Proposal p(x_t | x_{t-1}, y_t), weight p(y_t | x_{t-1}) for Gaussian transition noise and a linear
    likelihood (mocat/src/ssm/nonlinear_gaussian.py:134-281). Compiled for TimeHomogenousLinearGaussian
    (transition_function = A x), the model mocat's examples run it on. Initial weights are zero, so the
    log-likelihood returned estimates log p(y_1:T | y_0)."""
    name = "Optimal Non-linear Gaussian Particle Filter"
    kernel_model = 2

    def startup(self, ssm_scenario):
        """nonlinear_gaussian.py:154-184 in FP64."""
        M = _nlg_matrices(ssm_scenario)
        self.initial_kalman_gain, L, self.initial_conditioned_covariance_sqrt, icp = _initial_conditioning(M)
        self.initial_conditioned_precision_sqrt = L
        self.initial_conditioned_precision_det = np.linalg.det(icp)
        hq = M["H"] @ M["LQ"]
        W = np.linalg.inv(hq @ hq.T + M["LR"] @ M["LR"].T)
        self.proposal_kalman_gain = _kalman_gain(M["LQ"], M["H"], M["LR"], inv_mat=W)
        prop_cov = M["Q"] - self.proposal_kalman_gain @ M["H"] @ M["Q"]
        self.proposal_covariance_sqrt = _sym_chol(prop_cov)
        self.proposal_precision_sqrt = np.linalg.inv(self.proposal_covariance_sqrt)
        self.proposal_precision_det = 1 / np.linalg.det(prop_cov)
        self.weight_precision_sqrt = _sym_chol(W)
        self.weight_precision_det = np.linalg.det(W)

    def kernel_params(self, ssm_scenario):
        """[m0, H, K0, C0, A, K, S, WC, nlzw] for the fused filter kernel (model 2)."""
        self.startup(ssm_scenario)
        M = _nlg_matrices(ssm_scenario)
        dy = M["H"].shape[0]
        nlzw = (dy * LOG2PI - np.log(self.weight_precision_det)) / 2
        return np.concatenate([M["m0"].ravel(), M["H"].ravel(), self.initial_kalman_gain.ravel(),
                               self.initial_conditioned_covariance_sqrt.ravel(), M["A"].ravel(),
                               self.proposal_kalman_gain.ravel(), self.proposal_covariance_sqrt.ravel(),
                               self.weight_precision_sqrt.ravel(), [nlzw]]).astype(np.float32)


class EnsembleKalmanFilter(ParticleFilter):
    """Ensemble Kalman filter with perturbed observations (mocat/src/ssm/nonlinear_gaussian.py:284-350):
    forecast mx = f(x) + Q noise, gain from the forecast ensemble's covariance (ddof 1), analysis
    x = mx + K (y - H mx - R noise); log-weights stay zero. Compiled for TimeHomogenousLinearGaussian
    (transition_function = A x). run_particle_filter_for_marginals returns its ensembles and adds
    `gain` (T - 1, d, dy)."""
    name = "Ensemble Kalman Filter"

    def startup(self, ssm_scenario):
        """nonlinear_gaussian.py:297-310 in FP64."""
        M = _nlg_matrices(ssm_scenario)
        self.initial_kalman_gain, _, self.initial_conditioned_covariance_sqrt, _ = _initial_conditioning(M)

    def kernel_params(self, ssm_scenario):
        """[m0, H, K0, C0, A, LQ, LR] for the EnKF kernels."""
        self.startup(ssm_scenario)
        M = _nlg_matrices(ssm_scenario)
        return np.concatenate([M["m0"].ravel(), M["H"].ravel(), self.initial_kalman_gain.ravel(),
                               self.initial_conditioned_covariance_sqrt.ravel(), M["A"].ravel(), M["LQ"].ravel(),
                               M["LR"].ravel()]).astype(np.float32)


def _filter_kind(particle_filter):
    if particle_filter is None:
        return BootstrapFilter
    cls = particle_filter if isinstance(particle_filter, type) else type(particle_filter)
    for kind in (OptimalNonLinearGaussianParticleFilter, EnsembleKalmanFilter, BootstrapFilter):
        if issubclass(cls, kind):
            return kind
    raise NotImplementedError(f"{cls.__name__}: the compiled filters are BootstrapFilter, "
                              "OptimalNonLinearGaussianParticleFilter and EnsembleKalmanFilter")


def _run_enkf_hooks(model, enkf, yn, tn, random_key, n, store=True):
    """The EnKF at a (d, d_y) without fused kernels (up to 16): the key tree of mocat_enkf_run, the
    member updates by mocat_enkf_hook (initial ensemble; per step the forecast, then the analysis), and
    each step's gain K = P H^T (H P H^T + LR LR^T)^-1 from the forecast ensemble's covariance P (two
    passes in FP64: the mean, then the centred product as one GEMM), solved on the host in FP64."""
    d, dy = int(model.dim), int(model.dim_obs)
    dev = mr.default_device()
    T = int(yn.shape[0])
    params = torch.from_numpy(enkf.kernel_params(model)).to(dev)
    M = _nlg_matrices(model)
    H, LR = np.asarray(M["H"], np.float64), np.asarray(M["LR"], np.float64)
    y_t = torch.from_numpy(np.ascontiguousarray(yn, np.float32).reshape(T, dy)).to(dev)
    key = mr.as_key(random_key).reshape(2)
    rk, sub = mr.split(key, 2)

    def hook(phase, x_in, k, y, gain=None):
        out = torch.empty((n, d), dtype=torch.float32, device=dev)
        _lib.call("mocat_enkf_hook", phase, d, dy, _lib.ptr(params), _lib.ptr(x_in),
                  _lib.ptr(mr.as_key(k).reshape(2).contiguous()), int(n), _lib.ptr(y), _lib.ptr(gain), _lib.ptr(out),
                  _lib.stream_handle())
        return out

    x = hook(0, None, mr.split(sub, n)[0], y_t[0])
    values, gains = [x], []
    int_keys = mr.split(rk, T - 1) if T > 1 else None
    for t in range(1, T):
        k1, k2 = mr.split(mr.split(int_keys[t - 1], n)[0], 2)
        mx = hook(1, x, k1, None)
        c = mx.double()
        c = c - c.mean(0, keepdim=True)
        P = (c.T @ c / (n - 1)).cpu().numpy()
        B = H @ P
        K = np.linalg.solve(B @ H.T + LR @ LR.T, B).T
        Kt = torch.from_numpy(np.ascontiguousarray(K, np.float32)).to(dev)
        x = hook(2, mx, k2, y_t[t], Kt)
        gains.append(Kt)
        values.append(x)
    gain = torch.stack(gains) if gains else torch.empty((0, d, dy), dtype=torch.float32, device=dev)
    return cdict(value=torch.stack(values) if store else None,
                 log_weight=torch.zeros((T, n), dtype=torch.float32, device=dev),
                 t=torch.as_tensor(tn, device=dev), y=y_t, ess=torch.full((T,), float(n), device=dev), gain=gain,
                 final_value=x)


def _run_enkf(model, enkf, yn, tn, random_key, n, store=True):
    d, dy = int(model.dim), int(model.dim_obs)
    if n < 2:
        raise ValueError("the ensemble Kalman filter needs n >= 2 (its covariance divides by n - 1)")
    if n * max(d, dy) >= 2 ** 32:
        raise ValueError("the EnKF draws normal(key, (n, max(d, dy))): n * max(d, dy) must be < 2^32")
    if not _compiled_shape(0, d, dy):
        return _run_enkf_hooks(model, enkf, yn, tn, random_key, n, store)
    dev = mr.default_device()
    T = int(yn.shape[0])
    params = torch.from_numpy(enkf.kernel_params(model)).to(dev)
    y_t = torch.from_numpy(np.ascontiguousarray(yn, np.float32).reshape(T, dy)).to(dev)
    key = mr.as_key(random_key).reshape(2).contiguous().to(dev)
    nb = ctypes.c_int64(0)
    _lib.call("mocat_enkf_workspace_bytes", int(n), d, ctypes.byref(nb))
    ws = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    x = torch.empty((n, d), dtype=torch.float32, device=dev)
    value = torch.empty((T, n, d), dtype=torch.float32, device=dev) if store else None
    gain = torch.empty((max(T - 1, 0), d, dy), dtype=torch.float32, device=dev)
    _lib.call("mocat_enkf_run", d, dy, _lib.ptr(params), _lib.ptr(y_t), T, int(n), _lib.ptr(key), _lib.ptr(x),
              _lib.ptr(ws), int(ws.numel()), _lib.ptr(value), _lib.ptr(gain) if T > 1 else None, _lib.stream_handle())
    return cdict(value=value, log_weight=torch.zeros((T, n), dtype=torch.float32, device=dev),
                 t=torch.as_tensor(tn, device=dev), y=y_t, ess=torch.full((T,), float(n), device=dev), gain=gain,
                 final_value=x)


def resample_indices(weights: torch.Tensor, random_key, method: str = "systematic",
                     log_weights: bool = False) -> torch.Tensor:
    """Ancestor indices from non-negative weights (or log-weights, exponentiated as exp(w - max w)):
    exact fixed-point cumulative sums, uniforms from `random_key` (systematic: uniform(key);
    multinomial: uniform(key, (n,)))."""
    w = weights.to(torch.float32).contiguous()
    n = int(w.shape[0])
    key = mr.as_key(random_key).reshape(2).contiguous()
    ws = torch.empty(n, dtype=torch.int64, device=w.device)
    idx = torch.empty(n, dtype=torch.int32, device=w.device)
    _lib.call("mocat_resample", _lib.ptr(w), n, {"systematic": 0, "multinomial": 1}[method], int(log_weights),
              _lib.ptr(key),
              _lib.ptr(ws), _lib.ptr(idx), _lib.stream_handle())
    return idx


def categorical(random_key, log_weight: torch.Tensor, n_out: int = None) -> torch.Tensor:
    """random.categorical(key, log_weight, shape=(n_out,)) by Gumbel-max, as mocat's _resample."""
    lw = log_weight.to(torch.float32).contiguous()
    n = int(lw.shape[0])
    n_out = n if n_out is None else int(n_out)
    key = mr.as_key(random_key).reshape(2).contiguous()
    idx = torch.empty(n_out, dtype=torch.int32, device=lw.device)
    _lib.call("mocat_resample_categorical", _lib.ptr(lw), n, n_out, _lib.ptr(key), _lib.ptr(idx),
              _lib.stream_handle())
    return idx


def _key_np(k):
    if isinstance(k, torch.Tensor):
        return k.detach().cpu().numpy().astype(np.uint32).reshape(2)
    return np.asarray(k, np.uint32).reshape(2)


def _keys_np(keys):
    """Every filter's key as a row of one (F, 2) uint32 host array. Device keys cross to the host in one
    copy: a copy per key would synchronise with the device once per filter."""
    if (len(keys) > 1 and all(isinstance(k, torch.Tensor) and k.dtype in (torch.uint32, torch.int32) for k in keys)
            and len({k.device for k in keys}) == 1):
        rows = torch.stack([k.detach().reshape(2).view(torch.int32) for k in keys])
        return rows.cpu().numpy().view(np.uint32)
    return np.stack([_key_np(k) for k in keys]) if len(keys) else np.zeros((0, 2), np.uint32)


def _large_flags(ns, Ts, large_n):
    """Which filters take the chunked path: those with at least large_n particles (default 4096:
    above it one workgroup per filter is slower than a few grid-wide launches per time step)."""
    if large_n is None:
        large_n = 4096
    return [int(n) >= int(large_n) for n in ns]


def _compiled_shape(kind, d, dy) -> bool:
    """Whether the fused filters (mocat_pf_run / mocat_pf_run_large) have kernels for (kind, d, dy)."""
    c = _lib.ctypes.c_int(0)
    _lib.call("mocat_pf_compiled_shape", int(kind), int(d), int(dy), _lib.ctypes.byref(c))
    return bool(c.value)


def _run_lg_hooks(model, y, key, n, ess_threshold, resampling, store, initial=None, flags=0, optimal=None):
    """One bootstrap (or, with `optimal`, optimal-proposal) filter on the linear-Gaussian model at a
    (d, d_y) without fused kernels: the scan of run_particle_filter_for_marginals
    (mocat/src/ssm/filtering.py:255-324) with the fused filters' key tree, flags and log-likelihood
    recursion, as a host loop over the HIP kernels: the runtime-dimension hooks (initial sample and
    weight, proposal and weight; the optimal filter's noise from the first key of each split),
    mocat_gather_rows, the resampling entries (systematic / multinomial from the log-weights, or
    mocat's Gumbel-max) and the FP64 log-sum-exp. The resampling decision of each step reads the
    previous step's ESS on the host. Returns the fields of _launch's per-filter result."""
    from ..reduce import logsumexp_pair
    dev = mr.default_device()
    d, dy = int(model.dim), int(model.dim_obs)
    y = np.asarray(y, np.float32).reshape(-1, dy)
    T = y.shape[0]
    kid = 0 if optimal is None else OptimalNonLinearGaussianParticleFilter.kernel_model
    params = model.kernel_params() if optimal is None else optimal.kernel_params(model)
    pkeys = (lambda k: mr.split(k, n)) if optimal is None else (lambda k: mr.split(k, n)[:1])
    key = mr.as_key(key).reshape(2)
    resume = initial is not None
    sub = key if flags & PF_INIT_KEY else mr.split(key, 2)[1]
    rk = key if resume else mr.split(key, 2)[0]
    yt = torch.from_numpy(y).to(dev)
    if resume:
        x = torch.as_tensor(initial[0], device=dev).reshape(n, d).to(torch.float32).contiguous()
        lw = torch.as_tensor(initial[1], device=dev).reshape(n).to(torch.float32).contiguous()
    else:
        x, lw = _hook(kid, params, model, 0, None, yt[0], pkeys(sub), n)
    logn = float(np.log(n))
    lse, lse2 = (float(v) for v in logsumexp_pair(lw))
    ll = 0.0 if resume else lse - logn
    ess = [float(np.exp(2 * lse - lse2))]
    llt = [ll]
    ident = torch.arange(n, dtype=torch.int32, device=dev)
    xs, lws, ancs = [x], [lw], [ident]
    step_keys = None if flags & PF_STEP_KEY or T < 2 else mr.split(rk, T - 1)
    for t in range(1, T):
        ik = key if flags & PF_STEP_KEY else step_keys[t - 1]
        res = ess[-1] < ess_threshold * n
        if res:
            idx = categorical(ik, lw) if resampling == "categorical" else \
                resample_indices(lw, ik, resampling, log_weights=True)
            xp = torch.empty((n, d), dtype=torch.float32, device=dev)
            _lib.call("mocat_gather_rows", _lib.ptr(x), _lib.ptr(idx), n, d, _lib.ptr(xp), _lib.stream_handle())
        else:
            idx, xp = ident, x
        x, inc = _hook(kid, params, model, 1, xp, yt[t], pkeys(ik), n)
        lw = inc if res else lw + inc
        lse_prev = lse
        lse, lse2 = (float(v) for v in logsumexp_pair(lw))
        ll += lse - (logn if res else lse_prev)
        ess.append(float(np.exp(2 * lse - lse2)))
        llt.append(ll)
        if store:
            xs.append(x)
            lws.append(lw)
            ancs.append(idx)
    r = dict(log_likelihood=torch.tensor(ll, dtype=torch.float64, device=dev),
             ess=torch.tensor(ess, dtype=torch.float32, device=dev),
             log_likelihood_t=torch.tensor(llt, dtype=torch.float64, device=dev), final_value=x, final_log_weight=lw)
    if store:
        r.update(value=torch.stack(xs), log_weight=torch.stack(lws), ancestors=torch.stack(ancs))
    return r


def _launch(models, ys, keys, ns, ess_threshold, resampling, store, grad=False, large_n=None, particle_filter=None,
            initial=None, extra_flags=0, launch_mode=0, ts=None):
    """Runs len(models) filters (same model class and dims): filters below large_n particles in one
    persistent launch (a workgroup per filter), the others through the chunked per-step path
    (launch_mode 0: three stream launches per step; 1: those launches captured into one HIP graph;
    2: one persistent kernel for all steps)."""
    m0 = models[0]
    d, dy, kind = m0.dim, m0.dim_obs, m0.kernel_model
    for mdl in models:
        if (mdl.kernel_model, mdl.dim, mdl.dim_obs) != (kind, d, dy):
            raise ValueError("filters of one launch must share the model class and dimensions")
    optimal = _filter_kind(particle_filter) is OptimalNonLinearGaussianParticleFilter
    if optimal:
        opt = particle_filter if isinstance(particle_filter, ParticleFilter) else particle_filter()
        _nlg_matrices(m0)  # raises for models without a compiled optimal proposal
        kind = OptimalNonLinearGaussianParticleFilter.kernel_model
        if any(int(v) * d >= 2 ** 32 for v in ns):
            raise ValueError("the optimal filter draws normal(key, (n, d)): n * d must be < 2^32")
    if kind in (0, 2) and not _compiled_shape(kind, d, dy):
        # another linear-Gaussian shape: the runtime-dimension hooks, one filter at a time
        if grad:
            raise NotImplementedError("log-likelihood gradients are implemented for the 1-D stochastic-volatility model")
        return [_run_lg_hooks(models[i], ys[i], keys[i], int(ns[i]), ess_threshold, resampling, store,
                              None if initial is None else initial[i], extra_flags, opt if optimal else None)
                for i in range(len(models))]
    F = len(models)
    Ts = [int(np.asarray(y).shape[0]) for y in ys]
    order = sorted(range(F), key=lambda i: -(ns[i] * Ts[i]))  # longest first
    large = _large_flags(ns, Ts, large_n)
    desc = np.zeros(F, FILTER_DTYPE)
    keys_host = _keys_np(keys)
    params, yflat, scales = [], [], []
    poff = yoff = soff = ooff = toff = 0
    for slot, i in enumerate(order):
        p = opt.kernel_params(models[i]) if optimal else models[i].kernel_params()
        yi = np.asarray(ys[i], np.float32).reshape(Ts[i], dy)
        resume = initial is not None and initial[i] is not None
        desc[slot] = (ns[i], Ts[i], yoff, soff, ooff if store else -1, toff, keys_host[i], poff,
                      (PF_LARGE if large[i] else 0) | (PF_RESUME if resume else 0) | extra_flags)
        params.append(p)
        yflat.append(yi.ravel())
        scales.append(_scale_rows(models[i], None if ts is None else ts[i], Ts[i]))
        poff += p.size
        yoff += yi.size
        soff += ns[i]
        ooff += ns[i] * Ts[i]
        toff += Ts[i]
    dev = mr.default_device()
    f32 = dict(dtype=torch.float32, device=dev)
    desc_host = np.ascontiguousarray(desc)
    desc_t = torch.from_numpy(desc_host.view(np.uint8).copy()).to(dev)
    params_t = torch.from_numpy(np.concatenate(params)).to(dev)
    y_t = torch.from_numpy(np.concatenate(yflat).astype(np.float32)).to(dev)
    # per-row noise scales (row index = y_offset / dy + t), only when some filter has non-unit steps
    scale_t = None
    if any(sv is not None for sv in scales):
        scale_t = torch.from_numpy(np.concatenate([np.ones(len(yf) // dy, np.float32) if sv is None else sv
                                                   for sv, yf in zip(scales, yflat)])).to(dev)
    x0 = torch.empty(soff * d, **f32)
    x1 = torch.empty(soff * d, **f32)
    lw0 = torch.empty(soff, **f32)
    lw1 = torch.empty(soff, **f32)
    cs = torch.empty(soff, dtype=torch.int64, device=dev)
    anc = torch.empty(soff, dtype=torch.int32, device=dev)
    out_value = torch.empty(ooff * d, **f32) if store else None
    out_lw = torch.empty(ooff, **f32) if store else None
    out_anc = torch.empty(ooff, dtype=torch.int32, device=dev) if store else None
    out_ess = torch.empty(toff, **f32)
    out_llt = torch.empty(toff, dtype=torch.float64, device=dev)
    out_ll = torch.empty(F, dtype=torch.float64, device=dev)
    if grad and (kind != 1 or d != 1):
        raise NotImplementedError("log-likelihood gradients are implemented for the 1-D stochastic-volatility model")
    if grad and initial is not None and any(v is not None for v in initial):
        raise NotImplementedError("log-likelihood gradients start from the initial distribution, not from a sample")
    if initial is not None:
        # resumed filters read their t = 0 particles and log-weights from the state workspaces
        for slot, i in enumerate(order):
            if initial[i] is None:
                continue
            xi, li = initial[i]
            s0 = int(desc[slot]["state_offset"])
            x0[s0 * d:(s0 + ns[i]) * d] = torch.as_tensor(xi, device=dev).reshape(-1).to(torch.float32)
            lw0[s0:s0 + ns[i]] = torch.as_tensor(li, device=dev).reshape(-1).to(torch.float32)
    tan0 = torch.empty(soff * 5, **f32) if grad else None
    tan1 = torch.empty(soff * 5, **f32) if grad else None
    out_grad = torch.empty((F, 3), dtype=torch.float64, device=dev) if grad else None
    small = not all(large)
    cur = torch.cuda.current_stream()
    side = None
    if small:
        # the one-workgroup-per-filter kernel runs beside the chunked path's launches, except before the
        # persistent kernel (launch mode 2), whose grid barriers need all of its workgroups resident
        side = torch.cuda.Stream(device=dev) if any(large) and launch_mode != 2 else cur
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            _lib.call("mocat_pf_run", kind, d, dy, desc_host.ctypes.data_as(_lib.ctypes.c_void_p), _lib.ptr(desc_t),
                      F, _lib.ptr(params_t), _lib.ptr(y_t), _lib.ptr(scale_t),
                      _lib.ptr(x0), _lib.ptr(x1), _lib.ptr(lw0), _lib.ptr(lw1), _lib.ptr(cs), _lib.ptr(anc),
                      RESAMPLING[resampling], float(ess_threshold), _lib.ptr(out_value), _lib.ptr(out_lw),
                      _lib.ptr(out_anc), _lib.ptr(out_ess), _lib.ptr(out_llt), _lib.ptr(out_ll), _lib.ptr(tan0),
                      _lib.ptr(tan1), _lib.ptr(out_grad), _lib.stream_handle())
    if any(large):
        hp = desc_host.ctypes.data_as(_lib.ctypes.c_void_p)
        nb = _lib.ctypes.c_int64(0)
        _lib.call("mocat_pf_large_workspace_bytes", hp, F, _lib.ctypes.byref(nb))
        ws = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=dev)
        _lib.call("mocat_pf_run_large", kind, d, dy, hp, F, _lib.ptr(params_t), _lib.ptr(y_t), _lib.ptr(scale_t),
                  _lib.ptr(x0),
                  _lib.ptr(x1), _lib.ptr(lw0), _lib.ptr(lw1), _lib.ptr(cs), RESAMPLING[resampling],
                  float(ess_threshold), _lib.ptr(out_value), _lib.ptr(out_lw), _lib.ptr(out_anc), _lib.ptr(out_ess),
                  _lib.ptr(out_llt), _lib.ptr(out_ll), _lib.ptr(tan0), _lib.ptr(tan1), _lib.ptr(out_grad),
                  _lib.ptr(ws), int(ws.numel()), int(launch_mode), _lib.stream_handle())
    if side is not None and side is not cur:
        cur.wait_stream(side)
    res = [None] * F
    for slot, i in enumerate(order):
        r = dict(desc=desc[slot])
        n, T = ns[i], Ts[i]
        s0, o0, t0 = int(desc[slot]["state_offset"]), int(desc[slot]["out_offset"]), int(desc[slot]["t_offset"])
        r["log_likelihood"] = out_ll[slot]
        if grad:
            r["log_likelihood_grad"] = out_grad[slot]
        r["ess"] = out_ess[t0:t0 + T]
        r["log_likelihood_t"] = out_llt[t0:t0 + T]
        r["final_value"] = x0[s0 * d:(s0 + n) * d].reshape(n, d)
        r["final_log_weight"] = lw0[s0:s0 + n]
        if store:
            r["value"] = out_value[o0 * d:(o0 + n * T) * d].reshape(T, n, d)
            r["log_weight"] = out_lw[o0:o0 + n * T].reshape(T, n)
            r["ancestors"] = out_anc[o0:o0 + n * T].reshape(T, n)
        res[i] = r
    return res


def run_particle_filter_for_marginals(ssm_scenario, particle_filter, y, t, random_key, n: int = None,
                                      initial_sample: cdict = None, ess_threshold: float = 0.5,
                                      resampling: str = "categorical") -> cdict:
    kind = _filter_kind(particle_filter)
    if initial_sample is not None and kind is EnsembleKalmanFilter:
        raise NotImplementedError("the EnKF kernels start from the initial distribution (no initial_sample)")
    if resampling not in RESAMPLING:
        raise ValueError(f"resampling must be one of {list(RESAMPLING)}")
    if initial_sample is not None:
        n = int(initial_sample.value.shape[-2])
    if n is None:
        raise ValueError("n (number of particles) is required")
    yn = y.detach().cpu().numpy() if isinstance(y, torch.Tensor) else np.asarray(y)
    if yn.ndim == 1:
        yn = yn[:, None]
    tn = t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)
    if kind is EnsembleKalmanFilter:
        enkf = particle_filter if isinstance(particle_filter, ParticleFilter) else particle_filter()
        _nlg_matrices(ssm_scenario)
        return _run_enkf(ssm_scenario, enkf, yn, tn, random_key, int(n))
    if resampling == "categorical" and n * n >= 0xFFFFFFFF:
        raise ValueError("categorical resampling draws an (n, n) Gumbel matrix: n * n must be < 2^32 - 1")
    if initial_sample is not None:
        return _resume(ssm_scenario, particle_filter, yn, tn, random_key, int(n), initial_sample, ess_threshold,
                       resampling)
    r = _launch([ssm_scenario], [yn], [random_key], [int(n)], ess_threshold, resampling, True,
                particle_filter=particle_filter, ts=[tn])[0]
    dev = r["value"].device
    out = cdict(value=r["value"], log_weight=r["log_weight"], t=torch.as_tensor(tn, device=dev),
                y=torch.as_tensor(yn, dtype=torch.float32, device=dev), ess=r["ess"],
                log_likelihood=r["log_likelihood"], ancestors=r["ancestors"])
    return out


def _np2(y, dy=None):
    yn = y.detach().cpu().numpy() if isinstance(y, torch.Tensor) else np.asarray(y)
    return yn.reshape(-1, dy) if dy is not None else (yn[:, None] if yn.ndim == 1 else yn)


def initiate_particles(ssm_scenario, particle_filter, n: int, random_key, y=None, t: float = None) -> cdict:
    """mocat/src/ssm/filtering.py:173-193: particles from split(random_key, n) (the optimal filter:
    normal(split(random_key, n)[0], (n, d))), log-weights -likelihood(y) (zero without y, and always
    zero for the optimal filter). Returns cdict(value (1, n, d), log_weight (1, n), t (1,), y, ess (1,))."""
    kind = _filter_kind(particle_filter)
    if kind is EnsembleKalmanFilter:
        raise NotImplementedError("EnsembleKalmanFilter: use run_particle_filter_for_marginals")
    if kind is OptimalNonLinearGaussianParticleFilter and y is None:
        raise ValueError("the optimal filter's initial sample conditions on y")
    dy = int(ssm_scenario.dim_obs)
    yv = np.zeros((1, dy), np.float32) if y is None else _np2(y, dy)[:1]
    r = _launch([ssm_scenario], [yv], [random_key], [int(n)], 0.0, "systematic", True, particle_filter=particle_filter,
                extra_flags=PF_INIT_KEY)[0]
    dev = r["value"].device
    lw = r["log_weight"] if y is not None else torch.zeros((1, int(n)), dtype=torch.float32, device=dev)
    ess = r["ess"] if y is not None else torch.full((1,), float(n), device=dev)
    return cdict(value=r["value"], log_weight=lw,
                 t=torch.as_tensor([0.0 if t is None else float(t)], device=dev),
                 y=torch.as_tensor(yv, device=dev) if y is not None else None, ess=ess)


def resample_particles(particles: cdict, random_key, resample_full: bool = True) -> cdict:
    """mocat/src/ssm/filtering.py:200-215: ancestors random.categorical(random_key, log_weight[-1], (n,))
    (Gumbel-max, the categorical kernel); resample_full gathers every stored time point, else only the
    latest; the latest log-weights become zero."""
    out = particles.copy()
    n = int(particles.value.shape[-2])
    lw = particles.log_weight
    idx = categorical(random_key, lw if lw.dim() == 1 else lw[-1]).long()
    if lw.dim() == 1:
        out.value = particles.value[..., idx, :]
        out.log_weight = torch.zeros(n, dtype=lw.dtype, device=lw.device)
        return out
    if resample_full:
        out.value = particles.value[:, idx, :]
    else:
        out.value = particles.value.clone()
        out.value[-1] = particles.value[-1][idx]
    out.log_weight = lw.clone()
    out.log_weight[-1] = 0.0
    return out


def propagate_particle_filter(ssm_scenario, particle_filter, particles: cdict, y_new, t_new: float, random_key,
                              ess_threshold: float = 0.5, resample_full: bool = True) -> cdict:
    """mocat/src/ssm/filtering.py:218-253: resample (Gumbel-max with random_key) when the stored ESS
    particles.ess[-1] < ess_threshold * n, then one propagation step with the proposal keys
    split(random_key, n) (one resumed filter step that does not resample), appended to particles."""
    kind = _filter_kind(particle_filter)
    if kind is EnsembleKalmanFilter:
        raise NotImplementedError("EnsembleKalmanFilter: use run_particle_filter_for_marginals")
    n = int(particles.value.shape[1])
    if float(torch.as_tensor(particles.ess).reshape(-1)[-1]) < ess_threshold * n:
        particles = resample_particles(particles, random_key, resample_full)
    dy = int(ssm_scenario.dim_obs)
    yv = np.concatenate([np.zeros((1, dy), np.float32), _np2(y_new, dy)[:1].astype(np.float32)])
    t_last = getattr(particles, "t", None)
    tv = None if t_last is None else [float(torch.as_tensor(t_last).reshape(-1)[-1]), float(t_new)]
    r = _launch([ssm_scenario], [yv], [random_key], [n], 0.0, "systematic", True, particle_filter=particle_filter,
                initial=[(particles.value[-1], particles.log_weight[-1])], extra_flags=PF_STEP_KEY,
                ts=None if tv is None else [tv])[0]
    dev = r["value"].device
    out = particles.copy()
    out.value = torch.cat([particles.value, r["value"][1:]], 0)
    out.log_weight = torch.cat([particles.log_weight, r["log_weight"][1:]], 0)
    y_row = torch.as_tensor(yv[1:], device=dev)
    out.y = torch.cat([torch.as_tensor(particles.y, device=dev).reshape(-1, dy).to(y_row.dtype), y_row], 0) \
        if getattr(particles, "y", None) is not None else y_row
    out.t = torch.cat([torch.atleast_1d(torch.as_tensor(particles.t, device=dev)).to(torch.float32),
                       torch.as_tensor([float(t_new)], device=dev)])
    out.ess = torch.cat([torch.atleast_1d(torch.as_tensor(particles.ess, device=dev)).to(torch.float32), r["ess"][1:]])
    return out


def _resume(ssm_scenario, particle_filter, yn, tn, random_key, n, initial_sample, ess_threshold, resampling):
    """run_particle_filter_for_marginals from initial_sample (mocat/src/ssm/filtering.py:264-324): the
    scan starts from initial_sample[-1] with step keys split(random_key, len(y)) and every observation
    is a propagation step; the new steps are appended to the sample's arrays. The resampling decision of
    the first new step uses the ESS of the sample's last log-weights computed in FP64 (mocat reads its
    stored f32 ESS). log_likelihood estimates log p(y | the sample)."""
    if len(yn) == 0:
        return initial_sample.copy()
    dy = yn.shape[1]
    x_init = initial_sample.value[-1]
    lw_init = initial_sample.log_weight[-1]
    yfull = np.concatenate([np.zeros((1, dy), yn.dtype), yn])  # row 0 stands for the sample (not read)
    t_prev = getattr(initial_sample, "t", None)
    tfull = None if t_prev is None else np.concatenate([[float(torch.as_tensor(t_prev).reshape(-1)[-1])],
                                                        np.asarray(tn, np.float64).reshape(-1)])
    r = _launch([ssm_scenario], [yfull], [random_key], [n], ess_threshold, resampling, True,
                particle_filter=particle_filter, initial=[(x_init, lw_init)], ts=None if tfull is None else [tfull])[0]
    dev = r["value"].device
    out = initial_sample.copy()
    cat = lambda a, b: torch.cat([torch.as_tensor(a, device=dev).to(b.dtype), b], 0)
    out.value = cat(initial_sample.value, r["value"][1:])
    out.log_weight = cat(initial_sample.log_weight, r["log_weight"][1:])
    out.ess = cat(torch.atleast_1d(torch.as_tensor(initial_sample.ess, device=dev)), r["ess"][1:])
    y_new = torch.as_tensor(yn, dtype=torch.float32, device=dev)
    out.y = cat(initial_sample.y, y_new) if getattr(initial_sample, "y", None) is not None else y_new
    t_new = torch.as_tensor(tn, device=dev)
    out.t = cat(torch.atleast_1d(torch.as_tensor(initial_sample.t, device=dev)), t_new) \
        if getattr(initial_sample, "t", None) is not None else t_new
    if getattr(initial_sample, "ancestors", None) is not None:
        out.ancestors = cat(initial_sample.ancestors, r["ancestors"][1:])
    else:
        out.ancestors = r["ancestors"][1:]
    out.log_likelihood = r["log_likelihood"]
    return out


def run_particle_filters(ssm_scenarios, ys, random_keys, ns, ess_threshold: float = 0.5,
                         resampling: str = "systematic", store: bool = False, grad: bool = False,
                         large_n: int = None, particle_filter=None, schedule=None, ts=None):
    """Independent bootstrap filters of uneven sizes in one launch. Returns a list of cdicts with
    log_likelihood (FP64), ess (T,), final_value, final_log_weight (and value/log_weight/ancestors when
    store=True). grad=True (stochastic volatility): log_likelihood_grad = d log Z / d(phi, sigma, beta)
    of the fixed-resampling-index estimator. Filters with at least large_n particles (default 4096)
    run on the chunked path spread over the whole GPU; the results do not depend on the path.
    particle_filter: BootstrapFilter (default) or OptimalNonLinearGaussianParticleFilter.
    schedule: None (large_n as given, stream launches), a dict {"large_n": int, "launch": "stream" |
    "graph" | "persistent"} (or {"large_n": int, "graph": bool}), or "auto": the schedule autotuner (mocat_amd.ssm.schedule) times the candidate schedules on this
    batch the first time its shape is seen, records the fastest, and every later call of the same
    shape uses it.
    ts: optional list of observation-time arrays (one per filter, as ys); the stochastic-volatility
    transition noise then scales with sqrt(t_t - t_{t-1}) (None: unit steps)."""
    if _filter_kind(particle_filter) is EnsembleKalmanFilter:
        raise NotImplementedError("EnsembleKalmanFilter runs one ensemble per call: run_particle_filter_for_marginals")
    ys = [y.detach().cpu().numpy() if isinstance(y, torch.Tensor) else np.asarray(y) for y in ys]
    if resampling == "categorical" and any(int(v) * int(v) >= 0xFFFFFFFF for v in ns):
        raise ValueError("categorical resampling draws an (n, n) Gumbel matrix: n * n must be < 2^32 - 1")
    models, keys, ns = list(ssm_scenarios), list(random_keys), [int(v) for v in ns]
    mode = 0
    if schedule is not None:
        from . import schedule as sched
        if schedule == "auto":
            run = lambda c: _launch(models, ys, keys, ns, ess_threshold, resampling, store, grad, c["large_n"],
                                    particle_filter, launch_mode=sched.launch_mode(c), ts=ts)
            choice = sched.choose(sched.signature(models, ys, ns, resampling, grad, store, particle_filter), run)
        elif isinstance(schedule, dict):
            choice = schedule
        else:
            raise ValueError('schedule must be None, "auto" or a dict {"large_n": int, "launch": str}')
        large_n, mode = int(choice["large_n"]), sched.launch_mode(choice)
    res = _launch(models, ys, keys, ns, ess_threshold, resampling, store, grad, large_n, particle_filter,
                  launch_mode=mode, ts=ts)
    return [cdict(**{k: v for k, v in r.items() if k != "desc"}) for r in res]


class KalmanOutput(tuple):
    """(means, covariances) as mocat returns them, with the log-likelihood (and the same arrays) as
    attributes."""

    def __new__(cls, means, covs, log_likelihood):
        obj = super().__new__(cls, (means, covs))
        obj.value, obj.cov, obj.log_likelihood = means, covs, log_likelihood
        return obj


def _kalman_mats(lgssm):
    dev = mr.default_device()
    f = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64, device=dev)
    return (f(lgssm.transition_matrix), f(lgssm.transition_covariance), f(lgssm.likelihood_matrix),
            f(lgssm.likelihood_covariance), f(lgssm.initial_mean), f(lgssm.initial_covariance), dev)


def run_kalman_filter_for_marginals(lgssm, y, t=None):
    """Kalman filter means/covariances (mocat/src/ssm/linear_gaussian/kalman.py:16-57), returned as
    mocat's (means (T, d), covs (T, d, d)) tuple; `.log_likelihood` holds log p(y_0:T) (not in mocat).
    The t = 0 prior covariance is the initial covariance (mocat uses its square root, kalman.py:20,
    the same for the identity). FP64 torch on the GPU."""
    A, Q, H, R, m, P, dev = _kalman_mats(lgssm)
    yy = torch.as_tensor(np.asarray(y.detach().cpu() if isinstance(y, torch.Tensor) else y), dtype=torch.float64,
                         device=dev).reshape(len(y), -1)
    ll = torch.zeros((), dtype=torch.float64, device=dev)
    means, covs = [], []
    for k in range(yy.shape[0]):
        if k > 0:
            m = A @ m
            P = A @ P @ A.T + Q
        S = H @ P @ H.T + R
        r = yy[k] - H @ m
        L = torch.linalg.cholesky(S)
        sol = torch.cholesky_solve(r[:, None], L)[:, 0]
        ll = ll - 0.5 * (r @ sol + 2 * torch.log(torch.diagonal(L)).sum() + r.shape[0] * np.log(2 * np.pi))
        K = P @ H.T @ torch.cholesky_inverse(L)
        m = m + K @ r
        P = P - K @ H @ P
        means.append(m)
        covs.append(P)
    return KalmanOutput(torch.stack(means), torch.stack(covs), ll)


def run_kalman_smoother_for_marginals_and_lag_ones(lgssm, y, t=None, filter_output=None):
    """Rauch-Tung-Striebel backward pass (mocat/src/ssm/linear_gaussian/kalman.py:60-98) in FP64: back
    gain J = P_t A^T (A P_t A^T + Q)^-1, m_t = f_t + J (m_t+1 - A f_t), P_t = F_t + J (P_t+1 - A F_t A^T - Q) J^T,
    and mocat's lag-one term written as it is there, f_t . m_t+1 + J (P_t+1 + (m_t+1 - f_t+1) . m_t+1)
    (dot products of the mean vectors, as jnp's @ on 1-D arrays evaluates them). Returns
    (means (T, d), covs (T, d, d), lag_ones (T-1, d, d))."""
    A, Q, _, _, _, _, dev = _kalman_mats(lgssm)
    f_mus, f_covs = filter_output if filter_output is not None else run_kalman_filter_for_marginals(lgssm, y, t)
    f_mus = torch.as_tensor(f_mus, dtype=torch.float64, device=dev)
    f_covs = torch.as_tensor(f_covs, dtype=torch.float64, device=dev)
    T = f_mus.shape[0]
    mu, cov = f_mus[-1], f_covs[-1]
    mus, covs, lags = [mu], [cov], []
    for i in range(T - 2, -1, -1):
        fm, fc = f_mus[i], f_covs[i]
        J = fc @ A.T @ torch.linalg.inv(A @ fc @ A.T + Q)
        lag = fm @ mu + J @ (cov + (mu - f_mus[i + 1]) @ mu)
        mu, cov = fm + J @ (mu - A @ fm), fc + J @ (cov - A @ fc @ A.T - Q) @ J.T
        mus.append(mu)
        covs.append(cov)
        lags.append(lag)
    d = f_mus.shape[1]
    lag_t = torch.stack(lags[::-1]) if lags else torch.zeros((0, d, d), dtype=torch.float64, device=dev)
    return torch.stack(mus[::-1]), torch.stack(covs[::-1]), lag_t


def run_kalman_smoother_for_marginals(lgssm, y, t=None, filter_output=None):
    """mocat/src/ssm/linear_gaussian/kalman.py:101-107: (means, covs) of the smoothing marginals."""
    mus, covs, _ = run_kalman_smoother_for_marginals_and_lag_ones(lgssm, y, t, filter_output)
    return mus, covs
