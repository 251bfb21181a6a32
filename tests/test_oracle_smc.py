"""Pins the SMC oracle: mocat's bisect tests and the closed-form Gaussian evidence."""
import numpy as np

from oracle import rng, smc


def test_regula_falsi_reference_tests():
    # mocat/src/tests/test_utils.py:178-194
    f = lambda x: x ** 2 - 10
    b, bounds, evals, it = smc.regula_falsi(f, 0.0, 1e2)
    assert min(abs(e) for e in evals) < 1e-3
    b, bounds, evals, it = smc.regula_falsi(f, -1e1, 0.0)
    assert min(abs(e) for e in evals) < 1e-3
    assert abs(abs(b) - np.sqrt(10)) < 1e-4


def test_gaussian_evidence_closed_form_by_quadrature():
    from scipy import integrate
    m, v, pv, b = 0.7, 0.5, 4.0, 0.3
    f = lambda x: np.exp(-0.5 * x * x / pv) / np.sqrt(2 * np.pi * pv) * np.exp(-b * 0.5 * (x - m) ** 2 / v)
    q, _ = integrate.quad(f, -60, 60)
    np.testing.assert_allclose(smc.gaussian_log_evidence([m], [v], pv, b), np.log(q), rtol=1e-10)


def test_smc_log_evidence_matches_closed_form():
    mean, var, pv = np.array([1.0, -0.5]), np.array([0.3, 1.5]), 4.0
    out = smc.run(smc.gaussian_pot_grad(mean, var, pv), lambda k: np.sqrt(pv) * rng.normal(k, (2,)), 2, 1024,
                  rng.PRNGKey(0), stepsize=0.3, mcmc_steps=2, retain=0.5, thr_resample=1.0)
    assert out["temperature"][-1] == 1.0
    want = smc.gaussian_log_evidence(mean, var, pv, out["temperature"])
    np.testing.assert_allclose(out["log_norm_constant"], want, atol=0.08)
