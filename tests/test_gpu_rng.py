"""GPU RNG vs the oracle: bits and uniforms bit-exact, normals within the erfinv tolerance."""
import numpy as np
import pytest
import torch

from oracle import rng as orng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mr():
    from mocat_amd import random as mr
    return mr


def _np(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("num", [1, 2, 3, 7, 4096])
def test_split_bit_exact(mr, num):
    k = mr.PRNGKey(0)
    np.testing.assert_array_equal(_np(mr.split(k, num)), orng.split(orng.PRNGKey(0), num))


@pytest.mark.parametrize("size", [1, 2, 5, 50, 1023, 65537])
def test_bits_uniform_bit_exact(mr, size):
    k = mr.split(mr.PRNGKey(42), 3)[1]
    ko = orng.split(orng.PRNGKey(42), 3)[1]
    np.testing.assert_array_equal(_np(mr.bits(k, size)), orng.bits(ko, size))
    np.testing.assert_array_equal(_np(mr.uniform(k, size)), orng.uniform(ko, size))
    np.testing.assert_array_equal(_np(mr.uniform(k, size, -3.0, 2.5)), orng.uniform(ko, size, -3.0, 2.5))


def test_batched_keys_bit_exact(mr):
    keys = mr.split(mr.PRNGKey(1), 300)
    kn = orng.split(orng.PRNGKey(1), 300)
    got = _np(mr.split(keys, 11))
    want = np.stack([orng.split(k, 11) for k in kn])
    np.testing.assert_array_equal(got, want)
    got = _np(mr.bits(keys, (4, 3)))
    want = np.stack([orng.bits(k, (4, 3)) for k in kn])
    np.testing.assert_array_equal(got, want)


def test_normal_pdf_values_and_oracle(mr):
    k = mr.PRNGKey(0)
    _, s1, s2 = mr.split(k, 3)
    np.testing.assert_allclose([_np(mr.normal(k)).item(), _np(mr.normal(s1)).item(), _np(mr.normal(s2)).item()],
                               [-0.20584235, 0.5781487, 0.85355157], rtol=5e-7)
    x = np.float32(-5) + np.float32(0.1) * _np(mr.normal(s1, (4, 2)))
    np.testing.assert_allclose(x, [[-5.1619368, -4.871614], [-5.1136875, -5.048856],
                                   [-5.0195227, -4.8458843], [-5.037027, -5.017855]], rtol=1e-7)
    keys = mr.split(mr.PRNGKey(9), 64)
    kn = orng.split(orng.PRNGKey(9), 64)
    got = _np(mr.normal(keys, 1001))
    want = np.stack([orng.normal_f32(kk, 1001) for kk in kn])
    # same algorithm, device log1pf vs numpy log1p: a few f32 ulps apart at most
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-7)
    exact = np.stack([orng.normal_exact(kk, 1001) for kk in kn])
    np.testing.assert_allclose(got, exact, rtol=1e-5, atol=1e-6)


def test_errors_are_raised(mr):
    from mocat_amd._lib import MocatKernelError
    with pytest.raises(MocatKernelError):
        mr.bits(mr.PRNGKey(0), (2 ** 32,))


@pytest.mark.parametrize("lo,hi", [(0, 4096), (5, 17), (0, 1), (3, 3), (-7, 1000003), (0, 2 ** 31 - 1), (0, 300)])
@pytest.mark.parametrize("size", [1, 7, 4096, 30001])
def test_randint_bit_exact(mr, lo, hi, size):
    k = mr.split(mr.PRNGKey(11), 2)[1]
    ko = orng.split(orng.PRNGKey(11), 2)[1]
    np.testing.assert_array_equal(_np(mr.randint(k, (size,), lo, hi)), orng.randint(ko, (size,), lo, hi))


def test_randint_batched_keys_and_choice(mr):
    keys = mr.split(mr.PRNGKey(3), 5)
    kn = orng.split(orng.PRNGKey(3), 5)
    got = _np(mr.randint(keys, (6, 7), 2, 99))
    assert got.shape == (5, 6, 7)
    for i in range(5):
        np.testing.assert_array_equal(got[i], orng.randint(kn[i], (6, 7), 2, 99))
    np.testing.assert_array_equal(_np(mr.choice(keys[0], 4096, (64, 10))), orng.choice(kn[0], 4096, (64, 10)))
    a = torch.arange(100, 150, device="cuda") * 3
    np.testing.assert_array_equal(_np(mr.choice(keys[1], a, (20,))), _np(a)[orng.choice(kn[1], 50, (20,))])
