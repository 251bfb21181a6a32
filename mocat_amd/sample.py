"""Sampler base class and the `run` driver, mirroring mocat/src/sample.py.

`run(scenario, sampler, n, random_key, initial_state=None, initial_extra=None, **kwargs)` returns
a cdict whose attributes are stacked over the n iterations (initial state first), plus `time`
(wall seconds of the sampling loop, device-synchronised) and `summary` (mocat/src/sample.py:110-148).

Extension (opt-in keyword, reduces to mocat at its default): `n_chains=C` runs C independent chains
in one batched launch; chain c starts from key split(random_key, C)[c] and follows mocat's
single-chain key tree. Stacked attributes then have shape (n, C, ...).
"""
from __future__ import annotations

import copy
from inspect import isclass

from .core import Scenario, cdict, static_cdict


class Sampler:
    name: str = "Sampler"
    max_iter: int = 10000

    def __init__(self, name: str = None, **kwargs):
        if name is not None:
            self.name = name
        if not hasattr(self, "parameters") or "parameters" not in self.__dict__:
            self.parameters = cdict()
        for key, value in kwargs.items():
            if hasattr(self, key):
                setattr(self, key, value)
            else:
                setattr(self.parameters, key, value)

    def __repr__(self):
        return f"mocat_amd.Sampler.{self.__class__.__name__}"

    def deepcopy(self) -> "Sampler":
        return copy.deepcopy(self)

    def startup(self, scenario: Scenario, n: int, initial_state, initial_extra: cdict, **kwargs):
        for key, value in kwargs.items():
            if hasattr(self, key):
                setattr(self, key, value)
            if hasattr(self, "parameters") and hasattr(self.parameters, key):
                setattr(self.parameters, key, value)
        if not isinstance(self.max_iter, int):
            raise AttributeError(self.__repr__() + " max_iter must be int")
        if not hasattr(initial_extra, "iter"):
            initial_extra.iter = 0
        if not hasattr(initial_extra, "parameters"):
            initial_extra.parameters = cdict()
        for key, value in self.parameters.__dict__.items():
            if not hasattr(initial_extra.parameters, key) or getattr(initial_extra.parameters, key) is None:
                setattr(initial_extra.parameters, key, value)
        return initial_state, initial_extra

    def update(self, scenario: Scenario, state: cdict, extra: cdict):
        raise NotImplementedError(f"{self.name} update not initiated")

    def termination_criterion(self, state: cdict, extra: cdict) -> bool:
        return extra.iter >= self.max_iter

    def clean_chain(self, scenario: Scenario, chain_state: cdict):
        return chain_state

    def summary(self, scenario: Scenario, initial_state: cdict, initial_extra: cdict) -> cdict:
        summ = static_cdict()
        if hasattr(self, "name"):
            summ.sampler = self.name
        if hasattr(scenario, "name"):
            summ.scenario = scenario.name
        if hasattr(self, "parameters"):
            summ.parameters = self.parameters
        if hasattr(self, "tuning"):
            summ.tuning = self.tuning
        return summ

    # samplers with a fused device loop implement this: returns the stacked chain (no time/summary)
    def run_fused(self, scenario, n, initial_state, initial_extra, n_chains=None, **kwargs):
        raise NotImplementedError(
            f"{self.__class__.__name__} on {scenario.__class__.__name__} has no fused HIP path")


def run(scenario: Scenario, sampler, n: int, random_key, initial_state: cdict = None,
        initial_extra: cdict = None, n_chains: int = None, **kwargs) -> cdict:
    import time
    import torch

    if isclass(sampler):
        sampler = sampler(**kwargs)
    sampler.n = n
    if initial_extra is None:
        initial_extra = cdict()
    if random_key is not None:
        from . import random as mr
        initial_extra.random_key = mr.as_key(random_key)
    chain, elapsed, summary = sampler.run_fused(scenario, n, initial_state, initial_extra, n_chains=n_chains,
                                                **kwargs)
    chain.time = elapsed
    chain.summary = summary
    return chain
