"""mocat_resample_noalloc (library-owned scratch, no allocation) against mocat_resample (caller
workspace), through the C ABI: identical indices on two calls in a row, the second smaller."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_resample_noalloc_matches_workspace_entry_point():
    from mocat_amd import _lib
    import mocat_amd.ssm  # noqa: F401  (registers the signatures)
    stream = _lib.stream_handle()
    rs = np.random.default_rng(11)
    key = torch.tensor([7, 123456789], dtype=torch.uint32, device="cuda")
    for n in (300001, 4099):  # the second call smaller than the first
        for from_log in (0, 1):
            w = rs.standard_normal(n).astype(np.float32)
            w = w if from_log else np.abs(w)
            wt = torch.tensor(w, device="cuda")
            for method in (0, 1):
                ws = torch.empty(n, dtype=torch.int64, device="cuda")
                want = torch.empty(n, dtype=torch.int32, device="cuda")
                got = torch.full((n,), -1, dtype=torch.int32, device="cuda")
                _lib.call("mocat_resample", _lib.ptr(wt), n, method, from_log, _lib.ptr(key), _lib.ptr(ws),
                          _lib.ptr(want), stream)
                _lib.call("mocat_resample_noalloc", _lib.ptr(wt), n, method, from_log, _lib.ptr(key),
                          _lib.ptr(got), stream)
                torch.cuda.synchronize()
                assert torch.equal(got, want), (n, from_log, method)
    # beyond the scratch capacity: a clear size error, nothing launched
    big = (1 << 20) + 1
    fn = _lib._bind(_lib.load(), "mocat_resample_noalloc")
    dummy = torch.zeros(1, device="cuda")
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = fn(_lib.ptr(dummy), big, 0, 0, _lib.ptr(key), _lib.ptr(out), stream)
    assert rc == -3
