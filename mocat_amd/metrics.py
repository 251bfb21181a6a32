"""Run-quality metrics (mocat/src/metrics.py).

autocorrelation / integrated_autocorrelation_time / ess_autocorrelation follow
mocat/src/metrics.py:36-66 (FFT autocorrelation of the centred trace zero-padded to twice the next
power of two >= max_lag_eval; IAT = 1 + 2 cumsum(acf[1:]) at the first lag m with m > 5 IAT(m)),
batched here over trailing chain axes: a (T,) or (T, C) trace gives a scalar or (C,) result.
They are measurement utilities (used to report ESS/s), computed with torch FFTs in FP64.

log_ess_log_weight / ess_log_weight follow mocat/src/metrics.py:69-78:
2 * logsumexp(w) - logsumexp(2 w), reduced in FP64 by the reduction kernel of libmocat_amd.
"""
from __future__ import annotations

import torch


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
