"""FP64 log-sum-exp / ESS reductions over the last axis (libmocat_amd: mocat_logsumexp)."""
from __future__ import annotations

import math

import torch

from . import _lib

_lib.register("mocat_logsumexp", [_lib.c_ptr, _lib.c_i64, _lib.c_i64, _lib.c_ptr, _lib.c_ptr, _lib.c_ptr])


def logsumexp_pair(w: torch.Tensor):
    """(logsumexp(w), logsumexp(2 w)) over the last axis, FP64 results of shape w.shape[:-1]."""
    w = w.to(torch.float32).contiguous()
    batch = w.shape[:-1]
    B = int(math.prod(batch)) if len(batch) else 1
    a = torch.empty(batch, dtype=torch.float64, device=w.device)
    b = torch.empty(batch, dtype=torch.float64, device=w.device)
    _lib.call("mocat_logsumexp", _lib.ptr(w), B, int(w.shape[-1]), _lib.ptr(a), _lib.ptr(b), _lib.stream_handle())
    return a, b


def logsumexp(w: torch.Tensor) -> torch.Tensor:
    return logsumexp_pair(w)[0]


def log_ess(w: torch.Tensor) -> torch.Tensor:
    """2 logsumexp(w) - logsumexp(2 w) (mocat/src/metrics.py:69-74), accumulated in FP64."""
    a, b = logsumexp_pair(w)
    return 2 * a - b
