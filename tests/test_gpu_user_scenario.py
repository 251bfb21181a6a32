"""User scenarios (per-sample torch potentials, as mocat's Scenario subclasses are written,
mocat/src/core.py:146-230) on the device samplers: the Underdamped iteration in pieces
(mocat_hmc_begin / mocat_hmc_kick / mocat_hmc_correct) around the scenario's potentials, tempered SMC
moves and SVGD steps. Checked against the fused scenarios on the same keys, the oracle's iteration
(keys bit-exact, acceptance probabilities and accepted states) and the closed-form Gaussian evidence."""
import math

import numpy as np
import pytest
import torch

from oracle import lr as olr
from oracle import mcmc as omcmc
from oracle import rng as orng
from oracle import smc as osmc

pytestmark = pytest.mark.gpu

LOG2PI = math.log(2 * math.pi)


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def _user_gaussian(m, mean, var, prior_variance):
    """mocat_amd.DiagonalGaussian's potentials written as a user scenario."""
    mean_t = torch.tensor(mean, dtype=torch.float32, device="cuda")
    lam = 1.0 / torch.tensor(var, dtype=torch.float32, device="cuda")
    prec, d = 1.0 / prior_variance, len(mean)

    class UserGaussian(m.Scenario):
        name = "user gaussian"
        dim = d

        def prior_potential(self, x, random_key=None):
            return 0.5 * prec * torch.sum(x * x) + 0.5 * d * (LOG2PI - math.log(prec))

        def likelihood_potential(self, x, random_key=None):
            return 0.5 * torch.sum(lam * (x - mean_t) ** 2)

        def prior_sample(self, random_key):
            return math.sqrt(prior_variance) * m.random.normal(random_key, (d,))

    return UserGaussian()


def _user_lr(m, X, y, prior_variance=10.0):
    Xt = torch.tensor(X, dtype=torch.float32, device="cuda")
    yt = torch.tensor(y, dtype=torch.float32, device="cuda")
    prec, d = 1.0 / prior_variance, X.shape[1]

    class UserLR(m.Scenario):
        name = "user logistic regression"
        dim = d

        def prior_potential(self, x, random_key=None):
            return 0.5 * prec * torch.sum(x * x) + 0.5 * d * (LOG2PI - math.log(prec))

        def likelihood_potential(self, x, random_key=None):
            z = Xt @ x
            return torch.sum(torch.nn.functional.softplus(z) - yt * z)

        def prior_sample(self, random_key):
            return math.sqrt(prior_variance) * m.random.normal(random_key, (d,))

    return UserLR()


def test_user_gaussian_hmc_follows_the_fused_chain(m):
    mean, var, pv = [1.0, -0.5, 0.2, 2.0, 0.0], [0.3, 1.5, 0.7, 0.2, 1.0], 4.0
    fused = m.DiagonalGaussian(np.array(mean), np.array(var), prior_variance=pv)
    user = _user_gaussian(m, mean, var, pv)
    runs = []
    # a stepsize inside the leapfrog's stability region (h < 2 / sqrt(max precision) = 0.87 here): outside
    # it the integrator amplifies rounding differences of the potentials exponentially within a trajectory
    for sc in (fused, user):
        runs.append(m.run(sc, m.Underdamped(stepsize=0.3, leapfrog_steps=3), 25, m.random.PRNGKey(3), n_chains=64,
                          correction=m.Metropolis()))
    a, b = runs
    # same key tree and update arithmetic; the potentials differ only by f32 summation order, so the
    # chains coincide until (if ever) a uniform falls between the two acceptance probabilities
    diff = (a.value - b.value).abs().amax(dim=2)
    same = diff.amax(0) < 1e-5
    assert same.float().mean().item() >= 0.95, same.float().mean().item()  # measured: all 64, bit-identical
    torch.testing.assert_close(a.alpha[:, same], b.alpha[:, same], rtol=1e-4, atol=1e-5)
    torch.testing.assert_close(a.potential[:, same], b.potential[:, same], rtol=1e-5, atol=1e-4)
    torch.testing.assert_close(a.stepsize[:, same], b.stepsize[:, same], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("n,d,C,L", [(300, 7, 96, 4), (2000, 50, 64, 10)])
def test_user_lr_hmc_matches_oracle(m, n, d, C, L):
    X, y, _ = olr.make_dataset(3, n, d)
    sc = _user_lr(m, X, y)
    s = m.Underdamped(leapfrog_steps=L)
    extra = m.cdict(random_key=m.random.PRNGKey(5))
    state, extra = s.startup(sc, 10, None, extra, n_chains=C, correction=m.RMMetropolis())
    ost, okeys = omcmc.startup(orng.PRNGKey(5), C, d, lambda k: np.sqrt(10.0) * orng.normal(k, (d,)),
                               omcmc.lr_pot_grad_fn(X, y, 0.1))
    np.testing.assert_array_equal(extra.random_key.cpu().numpy(), okeys)
    np.testing.assert_allclose(state.value.cpu().numpy(), ost["value"], rtol=2e-6, atol=1e-6)
    s.iterate(sc, state, extra, 40)
    fields = ("value", "momenta", "potential", "grad_potential", "prior_potential", "likelihood_potential",
              "stepsize")
    host = {k: getattr(state, k).cpu().numpy().copy() for k in fields}
    keys0 = extra.random_key.cpu().numpy().copy()
    prev = state.value.cpu().numpy().copy()
    covered = 0
    for it in range(41, 44):
        s.update(sc, state, extra)
        o, okeys2, a, acc = omcmc.iterate(host, keys0, it, omcmc.lr_pot_grad_fn(X, y, 0.1), L, np.inf, "rm")
        np.testing.assert_array_equal(extra.random_key.cpu().numpy(), okeys2)
        ga, gv = state.alpha.cpu().numpy(), state.value.cpu().numpy()
        # f32 torch reductions of the user potential: relative error ~5e-6 of the potential's scale
        e_h = 5e-6 * (4.0 * np.abs(host["potential"]).astype(np.float64) + 10.0 * d)
        bound = 2e-5 + np.maximum(a, ga) * e_h
        assert np.all(np.abs(ga - a) <= bound), np.max(np.abs(ga - a) - bound)
        s2 = orng.split_batch(orng.split_batch(keys0, 2)[:, 0], L + 1)[:, 0]
        u = orng.uniform_batch(orng.split_batch(s2, 2)[:, 1], 1)[:, 0]
        g_acc = ~np.all(gv == prev, axis=1)
        clear = np.abs(u - a) > 2 * bound
        np.testing.assert_array_equal(g_acc[clear], acc[clear])
        both = clear & acc
        np.testing.assert_allclose(gv[both], o["value"][both], rtol=2e-4, atol=2e-4)
        covered += clear.sum()
        host = {k: getattr(state, k).cpu().numpy().copy() for k in host}
        keys0 = extra.random_key.cpu().numpy().copy()
        prev = gv.copy()
    assert covered >= 0.9 * 3 * C


def test_user_gaussian_smc_log_evidence_closed_form(m):
    mean, var, pv = np.array([1.0, -0.5]), np.array([0.3, 1.5]), 4.0
    sc = _user_gaussian(m, list(mean), list(var), pv)
    smp = m.MetropolisedSMCSampler(m.Underdamped(stepsize=0.3, leapfrog_steps=1), mcmc_steps=2,
                                   ess_threshold_retain=0.5, ess_threshold_resample=1.0, resampling="systematic")
    chain = m.run(sc, smp, 8192, m.random.PRNGKey(0))
    temps = chain.temperature.cpu().numpy()
    assert temps[-1] == 1.0 and np.all(np.diff(temps) > 0)
    want = osmc.gaussian_log_evidence(mean, var, pv, temps)
    np.testing.assert_allclose(chain.log_norm_constant.cpu().numpy(), want, atol=0.03)
    # the same run on the fused scenario visits the same first temperature (same prior draws)
    fused = m.run(m.DiagonalGaussian(mean, var, prior_variance=pv), smp, 8192, m.random.PRNGKey(0))
    np.testing.assert_allclose(temps[0], fused.temperature[0].item(), rtol=1e-5)


def test_user_mixture_svgd_follows_the_fused_steps(m):
    rng = np.random.default_rng(4)
    K, d, n = 3, 6, 512
    means = rng.standard_normal((K, d)) * 2
    var = rng.uniform(0.5, 2.0, (K, d))
    fused = m.GaussianMixture(means, var)
    mu, lam, logc = fused.means, fused.lam, fused.logc

    class UserMixture(m.Scenario):
        name = "user mixture"
        dim = d

        def prior_potential(self, x, random_key=None):
            return torch.sum(0.0 * x)

        def likelihood_potential(self, x, random_key=None):
            return -torch.logsumexp(logc - 0.5 * torch.sum(lam * (x - mu) ** 2, -1), 0)

    user = UserMixture()
    x0 = torch.tensor(rng.standard_normal((n, d)), dtype=torch.float32, device="cuda")
    out = []
    for sc in (fused, user):
        smp = m.SVGD(stepsize=0.1, max_iter=5, bandwidth_update="median")
        chain = m.run(sc, smp, n, m.random.PRNGKey(1), initial_state=m.cdict(value=x0.clone()), stack=("value",))
        out.append(chain.value)
    torch.testing.assert_close(out[0], out[1], rtol=1e-4, atol=1e-4)
