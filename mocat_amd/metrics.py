"""Run-quality metrics (mocat/src/metrics.py).

autocorrelation / integrated_autocorrelation_time / ess_autocorrelation follow
mocat/src/metrics.py:36-66 (FFT autocorrelation of the centred trace zero-padded to twice the next
power of two >= max_lag_eval; IAT = 1 + 2 cumsum(acf[1:]) at the first lag m with m > 5 IAT(m)),
batched here over trailing chain axes: a (T,) or (T, C) trace gives a scalar or (C,) result.
They are measurement utilities (used to report ESS/s), computed with torch FFTs in FP64.

log_ess_log_weight / ess_log_weight follow mocat/src/metrics.py:69-78:
2 * logsumexp(w) - logsumexp(2 w), reduced in FP64 by the reduction kernel of libmocat_amd.

ksd follows mocat/src/metrics.py:88-130 for the Gaussian and IMQ kernels over the full ensemble: the
weighted k_0 row sums come from the tiled pair kernel of libmocat_amd (FP32 coordinate sums, FP64
row sums), the final weighting and square root are FP64.
"""
from __future__ import annotations

import torch

from . import _lib as _l

_l.register("mocat_ksd_rows", [_l.c_ptr, _l.c_ptr, _l.c_ptr, _l.c_i64, _l.c_i64, _l.c_i32, _l.c_f32, _l.c_f32,
                               _l.c_f32, _l.c_i32, _l.c_ptr, _l.c_ptr])
_l.register("mocat_ksd_batch_rows", [_l.c_ptr, _l.c_ptr, _l.c_ptr, _l.c_i64, _l.c_i64, _l.c_i64, _l.c_ptr, _l.c_i32,
                                     _l.c_f32, _l.c_f32, _l.c_f32, _l.c_ptr, _l.c_ptr, _l.c_ptr])


def _next_pow_two(n: int) -> int:
    i = 1
    while i < n:
        i <<= 1
    return i


def autocorrelation(vals: torch.Tensor, max_lag_eval: int = 1000) -> torch.Tensor:
    v = vals.to(torch.float64)
    T = v.shape[0]
    max_lag_eval = min(max_lag_eval, T)
    n = _next_pow_two(max_lag_eval)
    x = v - v.mean(dim=0, keepdim=True)
    f = torch.fft.fft(x, n=2 * n, dim=0)
    acf = torch.fft.ifft(f * torch.conj(f), dim=0)[:max_lag_eval].real
    return acf / acf[0:1]


def integrated_autocorrelation_time(vals: torch.Tensor, max_lag_iat: int = None,
                                    max_lag_eval: int = 1000) -> torch.Tensor:
    ac = autocorrelation(vals, max_lag_eval)
    iats = 1 + 2 * torch.cumsum(ac[1:], dim=0)
    if max_lag_iat is None:
        c = 5
        m = torch.arange(1, iats.shape[0] + 1, device=iats.device, dtype=iats.dtype)
        if iats.dim() > 1:
            m = m.reshape(-1, *([1] * (iats.dim() - 1)))
        cond = (m - c * iats) > 0
        # argmax of a boolean: first True, or 0 when none
        idx = torch.argmax(cond.to(torch.int8), dim=0)
        return torch.gather(iats, 0, idx.unsqueeze(0)).squeeze(0)
    return iats[max_lag_iat]


def ess_autocorrelation(vals: torch.Tensor, max_lag_iat: int = None, max_lag_eval: int = 1000) -> torch.Tensor:
    """ESS of a (T,) or (T, C) trace (e.g. the potential), mocat/src/metrics.py:61-66."""
    if hasattr(vals, "potential") and not isinstance(vals, torch.Tensor):
        vals = vals.potential
    return vals.shape[0] / integrated_autocorrelation_time(vals, max_lag_iat, max_lag_eval)


def log_ess_log_weight(log_weight) -> torch.Tensor:
    from .reduce import log_ess
    if hasattr(log_weight, "log_weight") and not isinstance(log_weight, torch.Tensor):
        log_weight = log_weight.log_weight
    return log_ess(log_weight)


def ess_log_weight(log_weight) -> torch.Tensor:
    return torch.exp(log_ess_log_weight(log_weight))


_KSD_KERNELS = {"Gaussian": 0, "IMQ": 1}


def squared_jumping_distance(sample) -> torch.Tensor:
    """Mean squared increment of a chain along its first axis, per coordinate
    (mocat/src/metrics.py:81-85)."""
    vals = sample.value if hasattr(sample, "value") else sample
    vals = torch.as_tensor(vals)
    return ((vals[1:] - vals[:-1]) ** 2).mean(0)


def ksd(sample, kernel, grad_potential: torch.Tensor = None, log_weight: torch.Tensor = None,
        ensemble_batchsize: int = None, random_key=None, **kernel_params) -> float:
    """Kernel Stein discrepancy of a weighted ensemble (mocat/src/metrics.py:88-130).

    sample: (n, d) values or a cdict with .value (and .grad_potential when grad_potential is None).
    kernel: mocat_amd.kernels.Gaussian or IMQ; kernel_params override its parameters. As in mocat,
    k_0 takes the potential gradient where a Stein kernel takes the score. With ensemble_batchsize
    B, particle i's sum runs over row i of random.choice(random_key, n, (n, B)) (with replacement)
    and is normalised by the batch's weight sum, as mocat does."""
    from . import _lib
    vals = sample.value if hasattr(sample, "value") and not isinstance(sample, torch.Tensor) else sample
    if grad_potential is None and hasattr(sample, "grad_potential") and not isinstance(sample, torch.Tensor):
        grad_potential = sample.grad_potential
    elif grad_potential is None:
        raise TypeError("grad_potential not found")
    name = kernel.__class__.__name__
    if name not in _KSD_KERNELS:
        raise NotImplementedError(f"KSD is implemented for the Gaussian and IMQ kernels, not {name}")
    params = dict(kernel.parameters.__dict__)
    params.update(kernel_params)
    _lib.require_gpu()
    x = torch.as_tensor(vals, dtype=torch.float32, device="cuda")
    x = x.reshape(x.shape[0], -1).contiguous()
    g = torch.as_tensor(grad_potential, dtype=torch.float32, device="cuda").reshape(x.shape).contiguous()
    n, d = x.shape
    if log_weight is not None:
        w64 = torch.exp(torch.as_tensor(log_weight, device="cuda").to(torch.float64).reshape(n))
        w = w64.to(torch.float32).contiguous()
    else:
        w64, w = torch.ones(n, dtype=torch.float64, device="cuda"), None
    if ensemble_batchsize is not None:
        # metrics.py:101-104: batch_i = random.choice(random_key, n, (n, ensemble_batchsize))[i]
        if random_key is None:
            raise TypeError("ensemble_batchsize needs random_key")
        from . import random as mrandom
        batch = int(ensemble_batchsize)
        inds = mrandom.choice(random_key, n, (n, batch)).contiguous()
        row_k = torch.empty(n, dtype=torch.float64, device="cuda")
        row_w = torch.empty(n, dtype=torch.float64, device="cuda")
        _lib.call("mocat_ksd_batch_rows", _lib.ptr(x), _lib.ptr(g), _lib.ptr(w), n, d, batch, _lib.ptr(inds),
                  _KSD_KERNELS[name], float(params["bandwidth"]), float(params.get("c", 1.0)),
                  float(params.get("beta", -0.5)), _lib.ptr(row_k), _lib.ptr(row_w), _lib.stream_handle())
        return float(torch.sqrt((w64 * row_k / (w64.sum() * row_w)).sum()).item())
    tiles = -(-n // 64)
    js = max(1, min(tiles, 64))
    part = torch.empty((js, n), dtype=torch.float64, device="cuda")
    _lib.call("mocat_ksd_rows", _lib.ptr(x), _lib.ptr(g), _lib.ptr(w), n, d, _KSD_KERNELS[name],
              float(params["bandwidth"]), float(params.get("c", 1.0)), float(params.get("beta", -0.5)), js,
              _lib.ptr(part), _lib.stream_handle())
    s = w64.sum()
    return float(torch.sqrt((part.sum(0) * w64).sum() / (s * s)).item())
