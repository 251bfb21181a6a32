"""jax.random-compatible keys and draws, computed by HIP kernels on the GPU.

Keys are uint32 device tensors of shape (..., 2). Every function is batched over leading key
dimensions (the equivalent of ``vmap`` over keys): ``split(keys (K, 2), num)`` -> (K, num, 2),
``normal(keys (K, 2), shape)`` -> (K, *shape).

Bit layout follows the jax.random of the reference's era (reference call sites:
mocat/src/ssm/filtering.py:181,267,278,294; mocat/src/mcmc/standard_mcmc.py:94,120-122,129;
mocat/src/mcmc/metropolis.py:60-61): PRNGKey(s) = [0, s]; split and bits take the threefry2x32
outputs of a counter array cut into halves.
"""
from __future__ import annotations

import math
from typing import Sequence, Union

import numpy as np
import torch

from . import _lib

Shape = Union[int, Sequence[int]]


def _shape(shape: Shape) -> tuple:
    if isinstance(shape, int):
        return (shape,)
    return tuple(int(s) for s in shape)


def default_device() -> torch.device:
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def PRNGKey(seed: int, device=None) -> torch.Tensor:
    """Key for an integer seed: [seed >> 32, seed & 0xffffffff] with 32-bit seeds, i.e. [0, seed]."""
    seed = int(seed)
    if not (-(2 ** 31) <= seed < 2 ** 31):
        raise OverflowError("seed must fit in int32 (64-bit seeds are not enabled)")
    arr = np.array([0, seed & 0xFFFFFFFF], dtype=np.uint32)
    dev = default_device() if device is None else device
    return torch.from_numpy(arr).to(dev)


def as_key(key) -> torch.Tensor:
    if isinstance(key, torch.Tensor):
        if key.dtype != torch.uint32:
            if key.dtype in (torch.int32,):
                key = key.view(torch.uint32)
            else:
                raise TypeError(f"keys must be uint32, got {key.dtype}")
        if not key.is_cuda:
            key = key.to(default_device())
        if key.shape[-1] != 2:
            raise ValueError(f"keys must have a trailing dimension of 2, got {tuple(key.shape)}")
        return key.contiguous()
    arr = np.asarray(key, dtype=np.uint32)
    if arr.shape[-1] != 2:
        raise ValueError("keys must have a trailing dimension of 2")
    return torch.from_numpy(arr.copy()).to(default_device())


def _n_keys(key: torch.Tensor) -> int:
    return int(np.prod(key.shape[:-1], dtype=np.int64)) if key.dim() > 1 else 1


def split(key, num: int = 2) -> torch.Tensor:
    key = as_key(key)
    out = torch.empty((*key.shape[:-1], int(num), 2), dtype=torch.uint32, device=key.device)
    _lib.call("mocat_random_split", _lib.ptr(key), _n_keys(key), int(num), _lib.ptr(out), _lib.stream_handle())
    return out


def bits(key, shape: Shape = ()) -> torch.Tensor:
    key = as_key(key)
    shape = _shape(shape)
    size = math.prod(shape)
    out = torch.empty((*key.shape[:-1], *shape), dtype=torch.uint32, device=key.device)
    _lib.call("mocat_random_bits", _lib.ptr(key), _n_keys(key), size, _lib.ptr(out), _lib.stream_handle())
    return out


def uniform(key, shape: Shape = (), minval: float = 0.0, maxval: float = 1.0) -> torch.Tensor:
    key = as_key(key)
    shape = _shape(shape)
    size = math.prod(shape)
    out = torch.empty((*key.shape[:-1], *shape), dtype=torch.float32, device=key.device)
    _lib.call("mocat_random_uniform", _lib.ptr(key), _n_keys(key), size, float(np.float32(minval)),
              float(np.float32(maxval)), _lib.ptr(out), _lib.stream_handle())
    return out


def normal(key, shape: Shape = ()) -> torch.Tensor:
    key = as_key(key)
    shape = _shape(shape)
    size = math.prod(shape)
    out = torch.empty((*key.shape[:-1], *shape), dtype=torch.float32, device=key.device)
    _lib.call("mocat_random_normal", _lib.ptr(key), _n_keys(key), size, _lib.ptr(out), _lib.stream_handle())
    return out


def randint(key, shape: Shape = (), minval: int = 0, maxval: int = 1) -> torch.Tensor:
    """int32 draws in [minval, maxval) as jax.random.randint of the reference's era: two bit arrays
    from split(key), combined modulo the span with uint32 arithmetic."""
    key = as_key(key)
    shape = _shape(shape)
    size = math.prod(shape)
    out = torch.empty((*key.shape[:-1], *shape), dtype=torch.int32, device=key.device)
    _lib.call("mocat_random_randint", _lib.ptr(key), _n_keys(key), size, int(minval), int(maxval), _lib.ptr(out),
              _lib.stream_handle())
    return out


def choice(key, a, shape: Shape = (), replace: bool = True, p=None) -> torch.Tensor:
    """jax.random.choice with replacement and uniform probabilities (the form mocat calls,
    mocat/src/metrics.py:103): randint(key, shape, 0, n), then a[ind] when a is an array."""
    if not replace or p is not None:
        raise NotImplementedError("choice is implemented with replace=True and p=None")
    if isinstance(a, int):
        return randint(key, shape, 0, a)
    arr = torch.as_tensor(a)
    if arr.dim() != 1:
        raise ValueError("a must be an integer or one-dimensional")
    ind = randint(key, shape, 0, arr.shape[0])
    return arr.to(ind.device)[ind.long()]
