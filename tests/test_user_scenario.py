"""User scenarios written as in mocat (mocat/src/core.py:146-230): per-sample torch potentials, the
uniform-prior and potential-only rules of mocat's Scenario.__init__, and the derived gradients (per
sample and batched) against closed forms. CPU only (no kernels are called)."""
import math
import warnings

import pytest
import torch

from mocat_amd.core import Scenario


class Gauss(Scenario):
    name = "user gaussian"
    dim = 3

    def __init__(self, mean, lam, prec):
        self.mean, self.lam, self.prec = mean, lam, prec
        super().__init__()

    def prior_potential(self, x, random_key=None):
        return 0.5 * self.prec * torch.sum(x * x)

    def likelihood_potential(self, x, random_key=None):
        return 0.5 * torch.sum(self.lam * (x - self.mean) ** 2)


def _gauss():
    return Gauss(torch.tensor([1.0, -0.5, 0.2]), torch.tensor([2.0, 0.5, 1.5]), 0.1)


def test_derived_gradients_per_sample_and_batched():
    sc = _gauss()
    x = torch.randn(7, 4, 3, dtype=torch.float64)
    pp, lp, gp, gl = sc.evaluate(x, 0.3)
    assert pp.shape == (7, 4) and gl.shape == (7, 4, 3) and pp.dtype == torch.float32
    torch.testing.assert_close(gp, (0.1 * x).float())
    torch.testing.assert_close(gl, (sc.lam * (x - sc.mean)).float())
    torch.testing.assert_close(lp, (0.5 * torch.sum(sc.lam * (x - sc.mean) ** 2, -1)).float())
    # mocat's per-sample callables (value_and_grad of each potential) and the tempered combination
    v, g = sc.likelihood_potential_and_grad(x[0, 0])
    torch.testing.assert_close(g, sc.lam * (x[0, 0] - sc.mean))
    sc.temperature = 0.5
    v, g = sc.potential_and_grad(x[1, 2])
    torch.testing.assert_close(g, 0.1 * x[1, 2] + 0.5 * sc.lam * (x[1, 2] - sc.mean))
    torch.testing.assert_close(v, sc.prior_potential(x[1, 2]) + 0.5 * sc.likelihood_potential(x[1, 2]))
    torch.testing.assert_close(sc.potential(x[2, 1]), sc.prior_potential(x[2, 1]) + 0.5 * sc.likelihood_potential(x[2, 1]))
    torch.testing.assert_close(sc.potential(x)[2, 1].double(), sc.potential(x[2, 1]), rtol=1e-6, atol=1e-6)


def test_missing_prior_is_uniform_with_a_warning():
    class LikOnly(Scenario):
        name = "lik only"
        dim = 2

        def likelihood_potential(self, x, random_key=None):
            return torch.sum(torch.cosh(x))

    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        sc = LikOnly()
    assert any("assuming uniform" in str(e.message) for e in w)
    x = torch.randn(5, 2)
    pp, lp, gp, gl = sc.evaluate(x, 1.0)
    assert torch.all(pp == 0) and torch.all(gp == 0)
    torch.testing.assert_close(gl, torch.sinh(x))


def test_potential_only_becomes_the_likelihood():
    class PotOnly(Scenario):
        name = "potential only"
        dim = 2

        def potential(self, x, random_key=None):
            return torch.sum(x ** 4)

    sc = PotOnly()
    x = torch.randn(6, 2)
    pp, lp, gp, gl = sc.evaluate(x, 2.0)
    assert torch.all(pp == 0)
    torch.testing.assert_close(lp, torch.sum(x ** 4, -1))
    torch.testing.assert_close(gl, 4 * x ** 3)
    # potential is now the tempered one, as mocat's __init__ rebinds it
    sc.temperature = 2.0
    torch.testing.assert_close(sc.potential(x[0]), 2.0 * torch.sum(x[0] ** 4))


def test_keys_reach_the_potentials_per_sample():
    class Keyed(Scenario):
        name = "keyed"
        dim = 1

        def likelihood_potential(self, x, random_key=None):
            # a stochastic potential: the key's first word enters the value (not the gradient)
            return torch.sum(x * x) + random_key[0].to(x.dtype)

        def prior_potential(self, x, random_key=None):
            return torch.sum(0.0 * x)

    sc = Keyed()
    x = torch.zeros(4, 1)
    keys = torch.tensor([[1, 9], [2, 9], [3, 9], [4, 9]], dtype=torch.int64)
    _, lp, _, _ = sc.evaluate(x, 1.0, keys)
    torch.testing.assert_close(lp, torch.tensor([1.0, 2.0, 3.0, 4.0]))


def test_abstract_scenario_still_raises():
    sc = Scenario()
    with pytest.raises(AttributeError):
        sc.likelihood_potential(torch.zeros(2))
