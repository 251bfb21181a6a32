"""The particle-filter schedule records (mocat_amd.ssm.schedule): launch modes, signatures, the record
file, and the tuner's choice, on the CPU with a stand-in run function."""
import json

import numpy as np
import pytest

from mocat_amd.ssm import schedule as sched


def test_launch_mode_names_and_the_older_graph_flag():
    assert sched.launch_mode({"large_n": 4096, "launch": "stream"}) == 0
    assert sched.launch_mode({"large_n": 4096, "launch": "graph"}) == 1
    assert sched.launch_mode({"large_n": 4096, "launch": "persistent"}) == 2
    assert sched.launch_mode({"large_n": 4096, "graph": True}) == 1
    assert sched.launch_mode({"large_n": 4096, "graph": False}) == 0
    assert sched.launch_mode({"large_n": 4096}) == 0
    with pytest.raises(ValueError, match="launch must be one of"):
        sched.launch_mode({"large_n": 4096, "launch": "cooperative"})


class _Model:
    kernel_model, dim, dim_obs = 1, 1, 1


def test_signature_depends_on_the_sorted_shapes_only():
    ys = [np.zeros((5, 1)), np.zeros((9, 1))]
    a = sched.signature([_Model()] * 2, ys, [300, 700], "systematic", False, False, None)
    b = sched.signature([_Model()] * 2, ys[::-1], [700, 300], "systematic", False, False, None)
    c = sched.signature([_Model()] * 2, ys, [300, 701], "systematic", False, False, None)
    d = sched.signature([_Model()] * 2, ys, [300, 700], "systematic", True, False, None)
    assert a == b and a != c and a != d


def test_tuner_keeps_the_fastest_candidate_and_records_it(tmp_path, monkeypatch):
    monkeypatch.setenv("MOCAT_AMD_SCHEDULE_CACHE", str(tmp_path / "sched.json"))
    monkeypatch.setattr(sched, "_MEMO", {})
    monkeypatch.setattr(sched.torch.cuda, "synchronize", lambda: None)
    calls = []
    fake_ms = {(1, "stream"): 3.0, (1024, "graph"): 1.0}

    def run(c):
        calls.append(dict(c))
        clock[0] += fake_ms.get((c["large_n"], c["launch"]), 5.0) * 1e-3

    clock = [0.0]
    monkeypatch.setattr(sched.time, "perf_counter", lambda: clock[0])
    rec = sched.choose("sig", run)
    assert (rec["large_n"], rec["launch"]) == (1024, "graph")
    assert sched.launch_mode(rec) == 1
    n_cand = len(sched.CANDIDATE_LARGE_N) * len(sched.CANDIDATE_LAUNCH)
    assert len(rec["timings_ms"]) == n_cand and len(calls) == 3 * n_cand  # one untimed + two timed each
    assert json.loads((tmp_path / "sched.json").read_text())["sig"]["launch"] == "graph"
    # recorded: no more runs; a record without a launch entry (older format) is tuned again
    calls.clear()
    assert sched.choose("sig", run) == rec and not calls
    sched._MEMO["old"] = {"large_n": 2048, "graph": False, "timings_ms": {}}
    assert sched.choose("old", run)["launch"] == "graph" and calls


def test_filter_keys_cross_to_the_host_as_one_array():
    # the batch launch copies every filter's key in one transfer; uint32 words above 2^31 keep their
    # value, and keys of other integer types (or host arrays) take the per-key conversion
    import torch
    from mocat_amd.ssm.filtering import _key_np, _keys_np
    k32 = [torch.tensor(np.array([1, 0xFFFFFFFF], np.uint32)), torch.tensor(np.array([0x80000000, 4], np.uint32))]
    np.testing.assert_array_equal(_keys_np(k32), np.stack([_key_np(k) for k in k32]))
    assert _keys_np(k32).dtype == np.uint32
    k64 = [torch.tensor([1, 2], dtype=torch.int64), torch.tensor([3, 4], dtype=torch.int64)]
    np.testing.assert_array_equal(_keys_np(k64), np.array([[1, 2], [3, 4]], np.uint32))
    np.testing.assert_array_equal(_keys_np([np.array([5, 6], np.uint32)]), np.array([[5, 6]], np.uint32))
