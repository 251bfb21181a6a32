"""Schedule autotuner for batches of particle filters (run_particle_filters(..., schedule="auto")).

A batch of independent filters of uneven sizes is split between two launch structures: filters below
a particle-count threshold run one workgroup each in a single persistent launch; the others run on the
chunked path (1024-particle chunks spread over the GPU, three grid-wide phases per time step), either
as three launches per step queued one by one ("stream"), those launches captured into one HIP graph
("graph"), or one persistent kernel with grid barriers between the phases ("persistent"). Which split
and which launch mode is fastest depends on the batch (its size distribution and lengths), so the
tuner times every candidate schedule on the batch itself, keeps the fastest, and records it under the
batch's shape signature: in memory for the process and in a JSON file ($MOCAT_AMD_SCHEDULE_CACHE, else
~/.cache/mocat_amd/pf_schedule.json) for later processes. Results do not depend on the schedule (the
filters' key trees and arithmetic are the same on both paths), so tuning changes only the time.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

import numpy as np
import torch

CANDIDATE_LARGE_N = (1, 1024, 2048, 4096, 8192, 16384)  # 1: every filter on the chunked path
# the persistent kernel is not a default candidate: its grid barriers (an agent-scope L2 writeback and
# invalidate per workgroup) cost more than the kernel boundaries they replace on the batches measured
CANDIDATE_LAUNCH = ("stream", "graph")
LAUNCH_MODES = {"stream": 0, "graph": 1, "persistent": 2}

_MEMO: dict = {}


def launch_mode(choice: dict) -> int:
    """The chunked path's launch mode of a schedule: its "launch" entry, else the older "graph" flag."""
    if "launch" in choice:
        if choice["launch"] not in LAUNCH_MODES:
            raise ValueError(f"launch must be one of {sorted(LAUNCH_MODES)}")
        return LAUNCH_MODES[choice["launch"]]
    return 1 if choice.get("graph") else 0


def cache_path() -> Path:
    p = os.environ.get("MOCAT_AMD_SCHEDULE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "mocat_amd" / "pf_schedule.json"


def signature(models, ys, ns, resampling, grad, store, particle_filter) -> str:
    """Model class and dimensions, resampling, outputs, and the sorted (n, T) pairs of the batch."""
    m0 = models[0]
    shapes = np.asarray(sorted((int(n), int(np.asarray(y).shape[0])) for n, y in zip(ns, ys)), np.int64)
    h = hashlib.sha1(shapes.tobytes()).hexdigest()[:16]
    pf = "default" if particle_filter is None else getattr(particle_filter, "__name__", type(particle_filter).__name__)
    return f"{m0.kernel_model}-{m0.dim}-{m0.dim_obs}-{pf}-{resampling}-g{int(bool(grad))}-s{int(bool(store))}-{h}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(sig: str):
    if sig in _MEMO:
        return _MEMO[sig]
    rec = _load().get(sig)
    if rec is not None:
        _MEMO[sig] = rec
    return rec


def record(sig: str, choice: dict) -> None:
    _MEMO[sig] = choice
    path = cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        data = _load()
        data[sig] = choice
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs re-tuning in a later process


def tune(run, candidates=None, repeats: int = 2) -> dict:
    """Times run(schedule) for every candidate schedule (one untimed run each first, then the best of
    `repeats` timed runs, device-synchronised) and returns the fastest with all timings."""
    if candidates is None:
        candidates = [{"large_n": n, "launch": g} for n in CANDIDATE_LARGE_N for g in CANDIDATE_LAUNCH]
    timings = []
    for c in candidates:
        run(c)
        best = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(c)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        timings.append((best, c))
    timings.sort(key=lambda tc: tc[0])
    choice = dict(timings[0][1])
    choice["ms"] = timings[0][0] * 1e3
    choice["timings_ms"] = {f"large_n={c['large_n']},launch={c['launch']}": t * 1e3 for t, c in timings}
    return choice


def choose(sig: str, run) -> dict:
    """The recorded schedule for this signature, tuning (and recording) it on first use (records
    without a launch entry predate the persistent mode and are tuned again)."""
    rec = lookup(sig)
    if rec is None or "launch" not in rec:
        rec = tune(run)
        record(sig, rec)
    return rec
