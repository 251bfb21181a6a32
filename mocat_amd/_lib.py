"""Loader for libmocat_amd.so, the C-ABI kernel library, and thin call helpers.

The library is built in-tree by ``python -m mocat_amd.csrc.build``. torch is imported first so
that its HIP runtime (SONAME libamdhip64.so.7) is the one the kernel library binds to: device
pointers and ``hipStream_t`` handles from torch are then valid inside the library.

There is no fallback: if the library is missing, or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libmocat_amd.so"
ABI_VERSION = 5

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
c_ptr = ctypes.c_void_p

# name -> argtypes (every entry point returns int status)
SIGNATURES = {
    "mocat_abi_version": [],
    "mocat_random_split": [c_ptr, c_i64, c_i64, c_ptr, c_ptr],
    "mocat_random_bits": [c_ptr, c_i64, c_i64, c_ptr, c_ptr],
    "mocat_random_uniform": [c_ptr, c_i64, c_i64, c_f32, c_f32, c_ptr, c_ptr],
    "mocat_random_normal": [c_ptr, c_i64, c_i64, c_ptr, c_ptr],
    "mocat_random_randint": [c_ptr, c_i64, c_i64, c_i32, c_i32, c_ptr, c_ptr],
}

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP launch/runtime error", -3: "size limit exceeded"}


class MocatKernelError(RuntimeError):
    pass


_BOUND = {}


def register(name: str, argtypes):
    SIGNATURES[name] = argtypes
    _BOUND.pop(name, None)


def _bind(lib, name):
    """Resolve and type one entry point (on first use): a missing symbol fails that call only."""
    fn = _BOUND.get(name)
    if fn is None:
        try:
            fn = getattr(lib, name)
        except AttributeError as e:
            raise MocatKernelError(f"{name} is not exported by {LIB_PATH}: rebuild the library") from e
        fn.argtypes = SIGNATURES[name]
        fn.restype = ctypes.c_int
        _BOUND[name] = fn
    return fn


def load(path: os.PathLike | None = None):
    """Load (once) and return the ctypes handle. Raises if the library is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} not found: build the HIP kernels first (python -m mocat_amd.csrc.build). "
            "mocat_amd has no CPU fallback.")
    lib = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
    v = _bind(lib, "mocat_abi_version")()
    if v != ABI_VERSION:
        raise ImportError(f"libmocat_amd ABI version {v} != expected {ABI_VERSION}; rebuild the library")
    _LIB = lib
    return lib


def require_gpu(t: torch.Tensor | None = None):
    if not torch.cuda.is_available():
        raise MocatKernelError("mocat_amd kernels need a ROCm GPU (torch.cuda.is_available() is False)")
    if t is not None and not t.is_cuda:
        raise MocatKernelError("mocat_amd kernels take device tensors; got a host tensor")


def stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    require_gpu(t)
    if not t.is_contiguous():
        raise MocatKernelError("mocat_amd kernels take contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def call(name: str, *args):
    """Call an entry point (stream appended by the caller); raise on a non-zero status."""
    lib = load()
    rc = _bind(lib, name)(*args)
    if rc != 0:
        raise MocatKernelError(f"{name} failed with status {rc} ({STATUS.get(rc, 'unknown')})")
    return rc
