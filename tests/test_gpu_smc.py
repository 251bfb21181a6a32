"""GPU tempered SMC vs closed forms and the oracle; the leapfrog KAT through the fused kernel."""
import math

import numpy as np
import pytest
import torch

from oracle import lr as olr
from oracle import rng as orng
from oracle import smc as osmc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def test_leapfrog_reference_kat_on_gpu(m):
    # mocat/src/tests/test_utils.py:121-147 through the fused iteration kernel: U = 0.5|x|^2 (flat prior),
    # stepsize 0.1, 3 leapfrog steps from x = 0, p = 1, grad = [1, 2]. The kernel's `always` step
    # negates p (friction 0), so the chain starts from p = -1; accepted proposals store -p'.
    sc = m.DiagonalGaussian(np.zeros(2), np.ones(2), prior_variance=math.inf)
    s = m.Underdamped(stepsize=0.1, leapfrog_steps=3, friction=0.0)
    dev = "cuda"
    st = m.cdict(value=torch.zeros((1, 2), device=dev), momenta=-torch.ones((1, 2), device=dev),
                 grad_potential=torch.tensor([[1.0, 2.0]], device=dev), potential=torch.zeros(1, device=dev),
                 prior_potential=torch.zeros(1, device=dev), likelihood_potential=torch.zeros(1, device=dev),
                 stepsize=torch.full((1,), 0.1, device=dev), alpha=torch.ones(1, device=dev))
    s.correction = m.Uncorrected()
    ex = m.cdict(random_key=m.random.PRNGKey(0).reshape(1, 2), iter=0,
                 parameters=m.cdict(leapfrog_steps=3, friction=0.0, stepsize=0.1))
    s.iterate(sc, st, ex, 1)
    np.testing.assert_allclose(st.value[0].cpu().numpy(), [0.28120947, 0.266409], rtol=2e-7)
    np.testing.assert_allclose(st.grad_potential[0].cpu().numpy(), [0.28120947, 0.266409], rtol=2e-7)
    np.testing.assert_allclose(-st.momenta[0].cpu().numpy(), [0.9075344, 0.8597696], rtol=2e-7)


def test_smc_gaussian_log_evidence_closed_form(m):
    mean, var, pv = np.array([1.0, -0.5]), np.array([0.3, 1.5]), 4.0
    sc = m.DiagonalGaussian(mean, var, prior_variance=pv)
    smp = m.MetropolisedSMCSampler(m.Underdamped(stepsize=0.3, leapfrog_steps=1), mcmc_steps=2,
                                   ess_threshold_retain=0.5, ess_threshold_resample=1.0, resampling="systematic")
    chain = m.run(sc, smp, 8192, m.random.PRNGKey(0))
    temps = chain.temperature.cpu().numpy()
    assert temps[-1] == 1.0 and np.all(np.diff(temps) > 0)
    want = osmc.gaussian_log_evidence(mean, var, pv, temps)
    np.testing.assert_allclose(chain.log_norm_constant.cpu().numpy(), want, atol=0.03)
    # every visited temperature keeps ESS at the retain target (regula falsi converged)
    ess = chain.ess.cpu().numpy()[:-1]
    assert np.all(np.abs(ess / 8192 - 0.5) < 0.01)


def test_smc_first_temperature_matches_oracle(m):
    mean, var, pv = np.array([1.0, -0.5, 0.2]), np.array([0.3, 1.5, 0.7]), 4.0
    sc = m.DiagonalGaussian(mean, var, prior_variance=pv)
    smp = m.MetropolisedSMCSampler(m.Underdamped(stepsize=0.3), mcmc_steps=1, max_iter=1,
                                   ess_threshold_retain=0.5, ess_threshold_resample=1.0, resampling="systematic")
    chain = m.run(sc, smp, 1024, m.random.PRNGKey(4))
    o = osmc.run(osmc.gaussian_pot_grad(mean, var, pv), lambda k: np.sqrt(pv) * orng.normal(k, (3,)), 3, 1024,
                 orng.PRNGKey(4), stepsize=0.3, mcmc_steps=1, retain=0.5, thr_resample=1.0, max_iter=1)
    gt = chain.temperature.cpu().numpy()
    gz = chain.log_norm_constant.cpu().numpy()
    # startup: same prior draws and likelihoods -> same first temperature and log-normaliser
    np.testing.assert_allclose(gt[0], o["temperature"][0], rtol=1e-6)
    np.testing.assert_allclose(gz[0], o["log_norm_constant"][0], rtol=1e-5, atol=1e-6)
    # after one resample + move (f32 vs f64 weights may move a few ancestors)
    np.testing.assert_allclose(gt[1], o["temperature"][1], rtol=2e-2)
    np.testing.assert_allclose(gz[1], o["log_norm_constant"][1], atol=2e-2)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_smc_logistic_regression_vs_oracle(m, precision):
    X, y, _ = olr.make_dataset(5, 400, 6)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision=precision)
    smp = m.RMMetropolisedSMCSampler(m.Underdamped(stepsize=0.05), mcmc_steps=5, ess_threshold_retain=0.5,
                                     ess_threshold_resample=1.0, resampling="systematic")
    lz = []
    for s in range(4):
        chain = m.run(sc, smp, 4096, m.random.PRNGKey(s))
        assert chain.temperature[-1].item() == 1.0
        lz.append(chain.log_norm_constant[-1].item())
    pg = lambda x, b: olr.potential_and_grad(x, X, y, 0.1, b)
    o = [osmc.run(pg, lambda k: np.sqrt(10.0) * orng.normal(k, (6,)), 6, 1024, orng.PRNGKey(s), 0.05, 5, 1, 0.5,
                  1.0, "systematic", 1.0)["log_norm_constant"][-1] for s in range(3)]
    # same estimator, independent runs: log-evidences agree within Monte Carlo error
    assert abs(np.mean(lz) - np.mean(o)) < 4 * (np.std(o) + np.std(lz) + 0.05), (lz, o)


def test_smc_fixed_schedule(m):
    mean, var, pv = np.array([0.5]), np.array([0.2]), 2.0
    sc = m.DiagonalGaussian(mean, var, prior_variance=pv)
    sched = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    smp = m.MetropolisedSMCSampler(m.Underdamped(stepsize=0.2), mcmc_steps=2, temperature_schedule=sched,
                                   ess_threshold_resample=0.5, resampling="multinomial")
    chain = m.run(sc, smp, 4096, m.random.PRNGKey(1))
    np.testing.assert_allclose(chain.temperature.cpu().numpy(), sched[1:], rtol=1e-12)
    want = osmc.gaussian_log_evidence(mean, var, pv, sched[1:])
    np.testing.assert_allclose(chain.log_norm_constant.cpu().numpy(), want, atol=0.05)


def test_smc_logistic_regression_d100_vs_oracle(m):
    # the tempered-SMC move at d = 100: the d > 64 instantiation of the fused MALA kernel (config 3's
    # move, here on X 1000 x 100 so that the oracle finishes in seconds)
    n_data, d = 1000, 100
    X, y, _ = olr.make_dataset(9, n_data, d)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision="bf16x3")
    smp = m.RMMetropolisedSMCSampler(m.Underdamped(stepsize=0.02), mcmc_steps=5, ess_threshold_retain=0.5,
                                     ess_threshold_resample=1.0, resampling="systematic")
    lz, first = [], []
    for s in range(3):
        chain = m.run(sc, smp, 4096, m.random.PRNGKey(s))
        assert chain.temperature[-1].item() == 1.0
        assert np.isfinite(chain.log_norm_constant.cpu().numpy()).all()
        lz.append(chain.log_norm_constant[-1].item())
        first.append(chain.temperature[0].item())
    # startup: the prior draws of split(split(key)[1], n) and their FP64 likelihoods fix the first
    # temperature (regula falsi in FP64 on both sides); bf16x3 likelihood rounding moves it slightly
    for s in range(3):
        _, sub = orng.split(orng.PRNGKey(s), 2)
        x0 = math.sqrt(10.0) * orng.normal_batch(orng.split(sub, 4096), d).astype(np.float32)
        lik = olr.lik_potential(x0, X, y)
        b1 = osmc.next_temperature(np.zeros(4096), lik, 0.0, 4096.0, 0.5)[0]
        np.testing.assert_allclose(first[s], b1, rtol=1e-4)
    pg = lambda x, b: olr.potential_and_grad(x, X, y, 0.1, b)
    o = [osmc.run(pg, lambda k: np.sqrt(10.0) * orng.normal(k, (d,)), d, 1024, orng.PRNGKey(10 + s), 0.02, 5, 1,
                  0.5, 1.0, "systematic", 1.0)["log_norm_constant"][-1] for s in range(3)]
    # the same estimator: log-evidences agree within Monte Carlo error (the oracle runs 1024 particles)
    assert abs(np.mean(lz) - np.mean(o)) < 4 * (np.std(o) + np.std(lz) + 0.1), (lz, o)
