"""Kernels and bandwidth updates (mocat/src/kernels.py).

Gaussian(bandwidth): k(x, y) = exp(-|x - y|^2 / (2 h^2)) (kernels.py:82-116).
IMQ(c, beta, bandwidth): k(x, y) = (c + |x - y|^2 / (2 h^2))^beta (kernels.py:164-215).
median_bandwidth_update / mean_bandwidth_update: median / mean of all n^2 pairwise distances
(diagonal zeros included) divided by sqrt(2 log n) (kernels.py:220-229). Distances come from the
MFMA Gram kernel; the median is an exact radix select of the two middle order statistics.
"""
from __future__ import annotations

import math

import torch

from . import _lib
from .core import cdict

_P = _lib.c_ptr
_lib.register("mocat_svgd_workspace_bytes", [_lib.c_i64, _lib.c_i64, _lib.c_i32, _P])
_lib.register("mocat_pairwise_sqdist", [_P, _lib.c_i64, _lib.c_i64, _P, _P, _P])
_lib.register("mocat_select2", [_P, _lib.c_i64, _lib.c_i64, _lib.c_i64, _P, _P])
_lib.register("mocat_sqrt_mean", [_P, _lib.c_i64, _P, _lib.c_i32, _P])
_lib.register("mocat_pairwise_sqdist_upper", [_P, _lib.c_i64, _lib.c_i64, _P, _P, _P])
_lib.register("mocat_select2_upper", [_P, _lib.c_i64, _lib.c_i64, _lib.c_i64, _P, _P])
_lib.register("mocat_sqrt_sum_upper", [_P, _lib.c_i64, _P, _lib.c_i32, _P])


class Kernel:
    def __init__(self, **kwargs):
        if "parameters" not in self.__dict__:
            self.parameters = cdict()
        self.parameters.__dict__.update(kwargs)

    def __repr__(self):
        return f"mocat_amd.kernels.{self.__class__.__name__}"


class Gaussian(Kernel):
    def __init__(self, bandwidth: float = 1.0):
        super().__init__()
        self.parameters.bandwidth = bandwidth

    def __call__(self, x, y, bandwidth=None):
        h = self.parameters.bandwidth if bandwidth is None else bandwidth
        diff = (x - y) / h
        return torch.exp(-0.5 * torch.sum(diff * diff, dim=-1))


class IMQ(Kernel):
    """Inverse multiquadric k(x, y) = (c + |x - y|^2 / (2 h^2))^beta (mocat/src/kernels.py:164-215)."""

    def __init__(self, c: float = 1.0, beta: float = -0.5, bandwidth: float = 1.0):
        super().__init__()
        self.parameters.c = c
        self.parameters.beta = beta
        self.parameters.bandwidth = bandwidth

    def __call__(self, x, y, c=None, beta=None, bandwidth=None):
        c = self.parameters.c if c is None else c
        beta = self.parameters.beta if beta is None else beta
        h = self.parameters.bandwidth if bandwidth is None else bandwidth
        diff = (x - y) / h
        return torch.pow(c + 0.5 * torch.sum(diff * diff, dim=-1), beta)


def _workspace(n, d, js=1, device=None):
    b = _lib.ctypes.c_int64(0)
    _lib.call("mocat_svgd_workspace_bytes", int(n), int(d), int(js), _lib.ctypes.byref(b))
    return torch.empty(int(b.value), dtype=torch.uint8, device=device)


def pairwise_sqdist(vals: torch.Tensor) -> torch.Tensor:
    """All |x_i - x_j|^2 (n x n f32, diagonal exactly 0), MFMA Gram of the centred particles."""
    x = vals.to(torch.float32).contiguous()
    n, d = x.shape
    ws = _workspace(n, d, 1, x.device)
    d2 = torch.empty((n, n), dtype=torch.float32, device=x.device)
    _lib.call("mocat_pairwise_sqdist", _lib.ptr(x), n, d, _lib.ptr(ws), _lib.ptr(d2), _lib.stream_handle())
    return d2


def select_pair(v: torch.Tensor, rank0: int, rank1: int) -> torch.Tensor:
    """Exact rank0-th and rank1-th smallest of the non-negative floats in v (0-based), as f32."""
    v = v.to(torch.float32).contiguous().reshape(-1)
    ws = torch.empty(4 + 512, dtype=torch.int32, device=v.device)
    _lib.call("mocat_select2", _lib.ptr(v), int(v.numel()), int(rank0), int(rank1), _lib.ptr(ws),
              _lib.stream_handle())
    return ws[:2].view(torch.float32).clone()


def _sqdist_upper(vals: torch.Tensor) -> torch.Tensor:
    """Squared distances on and above the diagonal blocks of the n x n matrix (each pair computed once;
    the strict upper triangle, row < column, is what the selections read)."""
    x = vals.to(torch.float32).contiguous()
    n, d = x.shape
    ws = _workspace(n, d, 1, x.device)
    d2 = torch.empty((n, n), dtype=torch.float32, device=x.device)
    _lib.call("mocat_pairwise_sqdist_upper", _lib.ptr(x), n, d, _lib.ptr(ws), _lib.ptr(d2), _lib.stream_handle())
    return d2


def median_bandwidth_update(vals: torch.Tensor) -> torch.Tensor:
    """median(l2_distance_matrix(vals)) / sqrt(2 log n) (mocat/src/kernels.py:220-224), a device scalar.

    The n^2 distances are the n diagonal zeros and every pair twice, so the k-th smallest (k >= n) is
    the ((k - n) // 2)-th smallest pair distance: the selection runs over the n (n - 1) / 2 pairs."""
    n = vals.shape[0]
    m = n * n
    ranks = ((m - 1) // 2, m // 2)
    if n < 3:
        sel = select_pair(pairwise_sqdist(vals), *ranks)
    else:
        d2 = _sqdist_upper(vals)
        up = [max(k - n, 0) // 2 for k in ranks]
        ws = torch.empty(4 + 512, dtype=torch.int32, device=d2.device)
        _lib.call("mocat_select2_upper", _lib.ptr(d2), n, up[0], up[1], _lib.ptr(ws), _lib.stream_handle())
        sel = ws[:2].view(torch.float32).clone()  # both ranks are >= n when n >= 3
    med = 0.5 * (torch.sqrt(sel[0].double()) + torch.sqrt(sel[1].double()))
    return (med / math.sqrt(2 * math.log(n))).to(torch.float32)


def mean_bandwidth_update(vals: torch.Tensor) -> torch.Tensor:
    """mean(l2_distance_matrix(vals)) / sqrt(2 log n) (mocat/src/kernels.py:227-229), a device scalar:
    twice the FP64 sum of the pair distances over n^2."""
    n = vals.shape[0]
    parts = 1024
    if n < 2:
        return torch.full((), float("nan"), dtype=torch.float32, device=vals.device)
    d2 = _sqdist_upper(vals)
    partial = torch.empty(parts, dtype=torch.float64, device=d2.device)
    _lib.call("mocat_sqrt_sum_upper", _lib.ptr(d2), n, _lib.ptr(partial), parts, _lib.stream_handle())
    mean = 2.0 * partial.sum() / (n * n)
    return (mean / math.sqrt(2 * math.log(n))).to(torch.float32)
