"""cdict state container and the Scenario plug-in base, mirroring mocat/src/core.py.

`cdict` (mocat/src/core.py:20-84) is an attribute dictionary over torch device tensors:
indexing slices every non-scalar tensor attribute, `+` concatenates along axis 0 (chains stacked
over iterations), `save`/`load_cdict` persist it (tensors are written through torch.save with
weights_only-compatible content).

`Scenario` (mocat/src/core.py:124-261) keeps mocat's contract: prior_potential,
likelihood_potential, the tempered potential `prior + temperature * likelihood`, and the derived
`*_potential_and_grad` callables. Scenarios with a fused GPU kernel (scenarios.LogisticRegression,
scenarios.GaussianMixture) provide batched implementations of those callables directly: they take
(..., d) device tensors and return potentials (...,) and gradients (..., d).
"""
from __future__ import annotations

import copy
from pathlib import Path
from typing import Union

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


def save_cdict(in_cdict: cdict, path: Union[str, Path], overwrite: bool = False):
    """Persist a cdict (mocat/src/core.py:91-104) to a `.cdict` file (torch.save of host tensors)."""
    path = Path(path)
    if path.suffix != ".cdict":
        path = path.with_suffix(".cdict")
    path.parent.mkdir(parents=True, exist_ok=True)
    if path.exists():
        if overwrite:
            path.unlink()
        else:
            raise RuntimeError(f"File {path} already exists.")
    torch.save(_to_host(in_cdict), path)


def load_cdict(path: Union[str, Path], device=None) -> cdict:
    """Load a `.cdict` written by save_cdict (mocat/src/core.py:107-121); tensors go to `device`."""
    path = Path(path)
    if not path.is_file():
        raise ValueError(f"Not a file: {path}")
    if path.suffix != ".cdict":
        raise ValueError(f"Not a .cdict file: {path}")
    data = torch.load(path, weights_only=True)
    return _from_host(data, device)


def impl_checkable(func):
    func.not_implemented = True
    return func


def is_implemented(func):
    return not hasattr(func, "not_implemented")


class Scenario:
    """Target distribution: potential = prior_potential + temperature * likelihood_potential.

    Subclasses implement batched device versions of prior_potential / likelihood_potential and
    their `*_and_grad` variants over (..., d) tensors. mocat derives the gradients with jax.grad
    (mocat/src/core.py:204-230); here a scenario supplies them, typically through a fused kernel.
    """
    name: str = "Scenario"
    dim: int = None
    temperature: float = 1.0

    def __init__(self, name: str = None, **kwargs):
        if name is not None:
            self.name = name
        for key, value in kwargs.items():
            if hasattr(self, key):
                self.__dict__[key] = value

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
