"""cdict state container and the Scenario plug-in base, mirroring mocat/src/core.py.

`cdict` (mocat/src/core.py:20-84) is an attribute dictionary over torch device tensors:
indexing slices every non-scalar tensor attribute, `+` concatenates along axis 0 (chains stacked
over iterations), `save`/`load_cdict` persist it in mocat's `.cdict` format (a pickle of a
`mocat.src.core.cdict` whose arrays are numpy arrays, so mocat's `load_cdict` reads these files and
this module reads mocat's; loading goes through an unpickler that resolves only cdict and numpy
array classes).

`Scenario` (mocat/src/core.py:124-261) keeps mocat's contract: prior_potential,
likelihood_potential, the tempered potential `prior + temperature * likelihood`, and the derived
`*_potential_and_grad` callables. Scenarios with a fused GPU kernel (scenarios.LogisticRegression,
scenarios.GaussianMixture) provide batched implementations of those callables directly: they take
(..., d) device tensors and return potentials (...,) and gradients (..., d).
"""
from __future__ import annotations

import copy
import io
import pickle
from pathlib import Path
from typing import Union
from warnings import warn

import numpy as np
import torch


class cdict:
    def __init__(self, **kwargs):
        self.__dict__.update(kwargs)

    def copy(self) -> "cdict":
        return cdict(**self.__dict__)

    def deepcopy(self) -> "cdict":
        return copy.deepcopy(self)

    def __repr__(self):
        return f"mocat_amd.cdict({self.__dict__.__repr__()})"

    def save(self, path: Union[str, Path], overwrite: bool = False):
        save_cdict(self, path, overwrite)

    def __getitem__(self, item):
        if isinstance(item, str):
            return self.__dict__[item]
        out = self.copy()
        for key, attr in out.__dict__.items():
            if (isinstance(attr, torch.Tensor) and attr.dim() > 0) or (
                    isinstance(attr, cdict) and not isinstance(attr, static_cdict)):
                out.__setattr__(key, attr[item])
        return out

    def __add__(self, other: "cdict") -> "cdict":
        out = self.copy()
        if other is None:
            return out
        for key, attr in out.__dict__.items():
            if hasattr(other, key):
                o = other.__dict__[key]
                if isinstance(attr, torch.Tensor) or isinstance(o, torch.Tensor):
                    a = torch.atleast_1d(torch.as_tensor(attr))
                    b = torch.atleast_1d(torch.as_tensor(o, device=a.device))
                    out.__setattr__(key, torch.cat([a, b.to(a.dtype)], dim=0))
                elif ((isinstance(attr, cdict) and not isinstance(attr, static_cdict))
                      and isinstance(o, cdict) and not isinstance(o, static_cdict)) or key == "time":
                    out.__setattr__(key, attr + o)
        return out

    @property
    def is_empty(self):
        return self.__dict__ == {}

    def keys(self):
        return self.__dict__.keys()

    def __iter__(self):
        return self.__dict__.__iter__()


class static_cdict(cdict):
    pass


def _to_host(obj):
    if isinstance(obj, torch.Tensor):
        return obj.detach().cpu()
    if isinstance(obj, cdict):
        return {"__cdict__": type(obj).__name__, "items": {k: _to_host(v) for k, v in obj.__dict__.items()}}
    if isinstance(obj, (list, tuple)):
        return type(obj)(_to_host(v) for v in obj)
    if isinstance(obj, dict):
        return {k: _to_host(v) for k, v in obj.items()}
    return obj


def _from_host(obj, device):
    if isinstance(obj, torch.Tensor):
        return obj.to(device) if device is not None else obj
    if isinstance(obj, dict) and "__cdict__" in obj:
        cls = static_cdict if obj["__cdict__"] == "static_cdict" else cdict
        return cls(**{k: _from_host(v, device) for k, v in obj["items"].items()})
    if isinstance(obj, dict):
        return {k: _from_host(v, device) for k, v in obj.items()}
    return obj


# ---------------------------------------------------------------- mocat's .cdict format
# mocat writes pickle.dump(cdict) (mocat/src/core.py:91-104); its jax arrays pickle as numpy arrays and
# its load_cdict turns them back into jax arrays (:107-121). Files here hold the same objects: a
# `mocat.src.core.cdict` / `static_cdict` (module path and class name as mocat's) with numpy arrays for
# tensors, nested cdicts as such, and plain Python values.
_MOCAT_MODULE = "mocat.src.core"
_ZIP_MAGIC = b"PK\x03\x04"  # torch.save's zip container (files from earlier versions of this module)


class _Global:
    """Stands for the global `module.name` in a pickle (written by name, never called here)."""

    def __init__(self, module, name):
        self.module, self.name = module, name

    def __call__(self, *args, **kwargs):
        raise TypeError("a stand-in for a global in a pickle, not a callable")


class _MocatPickler(pickle._Pickler):
    """Pickles cdicts as mocat's classes and tensors as numpy arrays. Arrays name numpy's reconstructor
    by its numpy 1.x path (`numpy.core.multiarray`), which numpy 1.x and 2.x both resolve."""

    def reducer_override(self, obj):
        if isinstance(obj, cdict):
            cls = _Global(_MOCAT_MODULE, "static_cdict" if isinstance(obj, static_cdict) else "cdict")
            return cls, (), dict(obj.__dict__)
        if isinstance(obj, torch.Tensor):
            t = obj.detach().cpu()
            if t.dtype == torch.bfloat16:
                raise TypeError("a bfloat16 tensor has no numpy counterpart in mocat's .cdict format")
            func, args, *rest = t.numpy().__reduce_ex__(4)
            if getattr(func, "__name__", "") != "_reconstruct":
                raise TypeError(f"unexpected numpy reduction {func!r}")
            return (_Global("numpy.core.multiarray", "_reconstruct"), args, *rest)
        return NotImplemented

    def save(self, obj, save_persistent_id=True):
        if isinstance(obj, _Global):  # STACK_GLOBAL of its module and name
            self.save(obj.module)
            self.save(obj.name)
            self.write(pickle.STACK_GLOBAL)
            return
        super().save(obj, save_persistent_id)


# classes a .cdict file may name besides mocat's containers: what numpy arrays and scalars pickle as
_np_multiarray = getattr(np, "_core", None)
_np_multiarray = _np_multiarray.multiarray if _np_multiarray is not None else np.core.multiarray
_NUMPY_GLOBALS = {
    ("numpy._core.multiarray", "_reconstruct"): _np_multiarray._reconstruct,
    ("numpy.core.multiarray", "_reconstruct"): _np_multiarray._reconstruct,
    ("numpy._core.multiarray", "scalar"): _np_multiarray.scalar,
    ("numpy.core.multiarray", "scalar"): _np_multiarray.scalar,
    ("numpy", "ndarray"): np.ndarray,
    ("numpy", "dtype"): np.dtype,
    ("_codecs", "encode"): __import__("codecs").encode,
}


class _CdictUnpickler(pickle.Unpickler):
    """Resolves mocat's cdict classes to this module's and numpy's array classes; refuses every other
    global, so loading a file runs no code named by it."""

    def find_class(self, module, name):
        if module == _MOCAT_MODULE and name in ("cdict", "static_cdict"):
            return static_cdict if name == "static_cdict" else cdict
        if (module, name) in _NUMPY_GLOBALS:
            return _NUMPY_GLOBALS[(module, name)]
        raise pickle.UnpicklingError(f"a .cdict file may hold cdicts and numpy arrays only, not {module}.{name}")


def _to_torch(obj, device):
    if isinstance(obj, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(obj))
        return t.to(device) if device is not None else t
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, cdict):
        for k, v in list(obj.__dict__.items()):
            obj.__dict__[k] = _to_torch(v, device)
        return obj
    if isinstance(obj, (list, tuple)):
        return type(obj)(_to_torch(v, device) for v in obj)
    if isinstance(obj, dict):
        return {k: _to_torch(v, device) for k, v in obj.items()}
    return obj


def save_cdict(in_cdict: cdict, path: Union[str, Path], overwrite: bool = False):
    """Persist a cdict in mocat's .cdict format (mocat/src/core.py:91-104): the same suffix, overwrite and
    error rules, and a pickle that mocat's load_cdict reads (tensors are written as host numpy arrays)."""
    path = Path(path)
    if path.suffix != ".cdict":
        path = path.with_suffix(".cdict")
    path.parent.mkdir(parents=True, exist_ok=True)
    if path.exists():
        if overwrite:
            path.unlink()
        else:
            raise RuntimeError(f"File {path} already exists.")
    buf = io.BytesIO()
    _MocatPickler(buf, protocol=4).dump(in_cdict)
    path.write_bytes(buf.getvalue())


def load_cdict(path: Union[str, Path], device=None) -> cdict:
    """Load a .cdict file written by save_cdict or by mocat (mocat/src/core.py:107-121); numpy arrays
    become torch tensors on `device` (host tensors if None)."""
    path = Path(path)
    if not path.is_file():
        raise ValueError(f"Not a file: {path}")
    if path.suffix != ".cdict":
        raise ValueError(f"Not a .cdict file: {path}")
    raw = path.read_bytes()
    if raw[:4] == _ZIP_MAGIC:
        return _from_host(torch.load(io.BytesIO(raw), weights_only=True), device)
    data = _CdictUnpickler(io.BytesIO(raw)).load()
    if not isinstance(data, cdict):
        raise ValueError(f"{path} does not hold a cdict")
    return _to_torch(data, device)


def impl_checkable(func):
    func.not_implemented = True
    return func


def is_implemented(func):
    return not hasattr(func, "not_implemented")


def _zero_potential(x, random_key=None):
    return torch.zeros(x.shape[:-1], dtype=x.dtype, device=x.device)


def _grad_and_value(fn, x, random_key=None):
    """(grad, value) of a per-sample potential fn(x (d,), key) at x (..., d): torch.func.grad_and_value,
    vmapped over the leading dimensions (the torch counterpart of mocat's vmap(value_and_grad))."""
    from torch.func import grad_and_value, vmap
    if fn is _zero_potential:
        return torch.zeros_like(x), _zero_potential(x)
    if x.dim() <= 1:
        return grad_and_value(fn)(x, random_key)
    lead, d = x.shape[:-1], x.shape[-1]
    xf = x.reshape(-1, d)
    kf, kdim = None, None
    if random_key is not None:
        rk = torch.as_tensor(random_key)
        if rk.dim() > 1:
            kf, kdim = rk.reshape(-1, rk.shape[-1]), 0
        else:
            kf = rk
    g, v = vmap(grad_and_value(fn), in_dims=(0, kdim))(xf, kf)
    return g.reshape(*lead, d), v.reshape(lead)


class Scenario:
    """Target distribution: potential = prior_potential + temperature * likelihood_potential.

    Two kinds of scenario share this interface. Scenarios with a fused kernel
    (scenarios.LogisticRegression, DiagonalGaussian, GaussianMixture) implement batched device
    versions of the potentials and their `*_and_grad` variants over (..., d) tensors. A user scenario
    is written as in mocat (mocat/src/core.py:146-230): subclass Scenario and implement
    `likelihood_potential(x, random_key=None)` (or `potential`) and optionally `prior_potential`, each
    with torch operations on one sample x of shape (d,). As mocat's __init__ does, a missing prior
    becomes uniform (with a warning) and a lone `potential` becomes the likelihood; the gradients are
    derived automatically (torch.func.grad_and_value where mocat uses jax's value_and_grad) and every
    derived function also accepts a batch (..., d), evaluated with torch.func.vmap. The samplers run
    such scenarios with their device kernels and call the potentials between launches.
    """
    name: str = "Scenario"
    dim: int = None
    temperature: float = 1.0
    _fused_kernel = False  # set by the scenarios with a fused kernel

    def __init__(self, name: str = None, **kwargs):
        if name is not None:
            self.name = name
        for key, value in kwargs.items():
            if hasattr(self, key):
                self.__dict__[key] = value
        if not self._fused_kernel:
            self._derive_user_potentials()

    def _derive_user_potentials(self):
        """mocat's Scenario.__init__ / init_grad (mocat/src/core.py:161-230) for a user scenario."""
        has_lik = is_implemented(self.likelihood_potential)
        has_pot = type(self).potential is not Scenario.potential
        if not (has_lik or has_pot):
            return
        d = self.__dict__
        if not has_lik:
            if not is_implemented(self.prior_potential):
                d["prior_potential"] = _zero_potential
            d["likelihood_potential"] = self.potential
        if not is_implemented(self.prior_potential):
            warn(f"{self.name} prior_potential not initiated, assuming uniform")
            d["prior_potential"] = _zero_potential
        prior, lik = self.prior_potential, self.likelihood_potential
        d["tempered_potential"] = lambda x, temperature, random_key=None: \
            prior(x, random_key) + temperature * lik(x, random_key) if x.dim() <= 1 else \
            self.evaluate(x, temperature, random_key, grads=False)[0]
        d["potential"] = lambda x, random_key=None: self.tempered_potential(x, self.temperature, random_key)

        def pair(fn):
            def value_and_grad(x, random_key=None):
                g, v = _grad_and_value(fn, x, random_key)
                return v, g
            return value_and_grad

        d["prior_potential_and_grad"] = pair(prior)
        d["likelihood_potential_and_grad"] = pair(lik)

    def evaluate(self, x, temperature: float = 1.0, random_key=None, grads: bool = True):
        """(prior_potential, likelihood_potential, grad_prior, grad_likelihood) of a user scenario at
        x (..., d) in float32 on x's device (the batched form the samplers consume; the prior key is
        passed to both potentials, as mocat's samplers do). grads=False: tempered potential only."""
        if self._fused_kernel:
            raise NotImplementedError("fused scenarios define evaluate")
        x = torch.as_tensor(x)
        if not grads:
            from torch.func import vmap
            prior, lik = self.prior_potential, self.likelihood_potential
            xf = x.reshape(-1, x.shape[-1])
            kdim = 0 if (random_key is not None and torch.as_tensor(random_key).dim() > 1) else None
            kf = None if random_key is None else torch.as_tensor(random_key).reshape(
                (-1, 2) if kdim == 0 else (2,))
            f = lambda xx, kk: prior(xx, kk) + temperature * lik(xx, kk)
            return (vmap(f, in_dims=(0, kdim))(xf, kf).reshape(x.shape[:-1]),)
        gp, pp = _grad_and_value(self.prior_potential, x, random_key)
        gl, lp = _grad_and_value(self.likelihood_potential, x, random_key)
        f32 = lambda t: t.to(torch.float32).contiguous()
        return f32(pp), f32(lp), f32(gp), f32(gl)

    def __repr__(self):
        return f"mocat_amd.Scenario.{self.__class__.__name__}"

    # -- potentials (override) ------------------------------------------------------------------
    @impl_checkable
    def prior_potential(self, x, random_key=None):
        raise AttributeError(f"{self.name} prior_potential not initiated")

    @impl_checkable
    def likelihood_potential(self, x, random_key=None):
        raise AttributeError(f"{self.name} likelihood_potential not initiated")

    @impl_checkable
    def prior_sample(self, random_key):
        raise AttributeError(f"{self.name} prior_sample not initiated")

    @impl_checkable
    def prior_potential_and_grad(self, x, random_key=None):
        raise AttributeError(f"{self.name} prior_potential_and_grad not initiated")

    @impl_checkable
    def likelihood_potential_and_grad(self, x, random_key=None):
        raise AttributeError(f"{self.name} likelihood_potential_and_grad not initiated")

    # -- derived (mocat/src/core.py:173-230) -------------------------------------------------------
    def tempered_potential(self, x, temperature, random_key=None):
        return self.prior_potential(x, random_key) + temperature * self.likelihood_potential(x, random_key)

    def potential(self, x, random_key=None):
        return self.tempered_potential(x, self.temperature, random_key)

    def tempered_potential_and_grad(self, x, temperature, random_key=None):
        pp, pg = self.prior_potential_and_grad(x, random_key)
        lp, lg = self.likelihood_potential_and_grad(x, random_key)
        return pp + temperature * lp, pg + temperature * lg

    def potential_and_grad(self, x, random_key=None):
        return self.tempered_potential_and_grad(x, self.temperature, random_key)

    def grad_potential(self, x, random_key=None):
        return self.potential_and_grad(x, random_key)[1]

    def grad_prior_potential(self, x, random_key=None):
        return self.prior_potential_and_grad(x, random_key)[1]

    def grad_likelihood_potential(self, x, random_key=None):
        return self.likelihood_potential_and_grad(x, random_key)[1]
