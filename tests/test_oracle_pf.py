"""Pins the particle-filter, resampling and Kalman oracles."""
import numpy as np
import pytest

from oracle import kalman, pf, rng


def _lgssm(seed=0, d=4):
    rs = np.random.default_rng(seed)
    A = 0.9 * np.eye(d) + np.where(np.eye(d) > 0, 0, rs.normal(0, 0.1, (d, d)))
    H = rs.normal(0, 0.5, (d, d))
    return pf.lgssm_params(np.zeros(d), np.eye(d), A, 0.1 * np.eye(d), H, 0.5 * np.eye(d))


def test_kalman_matches_bruteforce_joint_gaussian():
    p = _lgssm(1, 3)
    _, y = pf.simulate(p, 6, rng.PRNGKey(2))
    P0 = np.diag([1.0, 2.0, 0.5])  # non-identity P0: exercises the corrected t = 0 covariance
    Q, R = p["LQ"] @ p["LQ"].T, np.linalg.inv(p["SR"] @ p["SR"].T)
    ll, _, _ = kalman.loglik(p["m0"], P0, p["A"], Q, p["H"], R, y)
    bf = kalman.joint_loglik_bruteforce(p["m0"], P0, p["A"], Q, p["H"], R, y)
    assert abs(ll - bf) < 1e-9 * abs(bf)


def test_lgssm_likelihood_is_gaussian_logpdf():
    from scipy.stats import multivariate_normal
    p = _lgssm(3)
    x = np.random.default_rng(0).standard_normal((5, 4))
    y = np.random.default_rng(1).standard_normal(4)
    R = np.linalg.inv(p["SR"] @ p["SR"].T)
    want = -multivariate_normal(mean=np.zeros(4), cov=R).logpdf(y[None, :] - x @ p["H"].T)
    np.testing.assert_allclose(pf.lik_potential(p, x, y), want, rtol=1e-10)


def test_sv_likelihood_is_gaussian_logpdf():
    from scipy.stats import norm
    p = pf.sv_params(0.95, 0.2, 0.7)
    x = np.array([[-1.0], [0.3], [2.0]])
    want = -norm(0, 0.7 * np.exp(0.5 * x[:, 0])).logpdf(0.4)
    np.testing.assert_allclose(pf.lik_potential(p, x, np.array([0.4])), want, rtol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 7, 100, 1000])
def test_systematic_offspring_bounds_and_bruteforce(n):
    rs = np.random.default_rng(n)
    w = rs.uniform(0, 1, n).astype(np.float32) * (rs.uniform(size=n) > 0.3)
    if w.sum() == 0:
        w[0] = 1
    for u in (0.0, 0.37, 0.999):
        a = pf.systematic(w, u)
        assert np.all(np.diff(a) >= 0)
        q = pf.quantise(w).astype(np.float64)
        # brute force: count j with threshold in [C_{i-1}, C_i)
        C = np.cumsum(q)
        T = (np.arange(n) + u) * (C[-1] / n)
        bf = np.array([np.argmax(C > t) if np.any(C > t) else n - 1 for t in T])
        np.testing.assert_array_equal(a, bf)
        counts = np.bincount(a, minlength=n)
        expect = n * q / q.sum()
        assert np.all(counts >= np.floor(expect) - 1) and np.all(counts <= np.ceil(expect) + 1)
        assert np.all(counts[q == 0] == 0)


def test_multinomial_frequencies():
    w = np.array([0.1, 0.0, 0.5, 0.4], np.float32)
    a = pf.resample_indices(w, rng.PRNGKey(3), "multinomial")
    assert a.shape == (4,)
    a = pf.multinomial(w, rng.uniform(rng.PRNGKey(4), (20000,)))
    freq = np.bincount(a, minlength=4) / 20000
    np.testing.assert_allclose(freq, w / w.sum(), atol=0.015)


def test_quantise_is_exact_scaling():
    w = np.array([1.0, 0.5, 0.25, 3e-8, 0.0], np.float32)
    q = pf.quantise(w)
    # max weight maps to 2^(52 - bitlen(n) - 1): scaling by a power of two keeps ratios exact
    assert q[0] == 2 * q[1] == 4 * q[2]
    assert q[4] == 0


def test_pf_loglik_close_to_kalman_statistically():
    p = _lgssm(0, 2)
    _, y = pf.simulate(p, 30, rng.PRNGKey(0))
    Q, R = p["LQ"] @ p["LQ"].T, np.linalg.inv(p["SR"] @ p["SR"].T)
    ll, _, _ = kalman.loglik(p["m0"], p["L0"] @ p["L0"].T, p["A"], Q, p["H"], R, y)
    lls = [pf.run_pf(p, y, rng.PRNGKey(s), 512, 1.0, "systematic")["log_likelihood"] for s in range(12)]
    m, se = np.mean(lls), np.std(lls) / np.sqrt(len(lls))
    assert abs(m - ll) < 4 * se + 0.05, (m, se, ll)



@pytest.mark.parametrize("timed", [False, True])
def test_sv_tangent_gradient_matches_finite_differences(timed):
    # fixed-resampling-index estimator: ancestors and noise held fixed; log Z is then smooth in the
    # parameters and its gradient is checked by central differences of the same oracle filter (also
    # with irregular observation times, where the noise scales with sqrt(dt))
    phi, sigma, beta, n, T = 0.95, 0.25, 0.7, 200, 25
    p = pf.sv_params(phi, sigma, beta)
    _, y = pf.simulate(p, T, rng.PRNGKey(5))
    y = y[:, 0]
    tt = np.cumsum(np.random.default_rng(9).uniform(0.2, 3.0, T)) if timed else None
    base = pf.run_pf(p, y, rng.PRNGKey(6), n, 1.0, "systematic", t=tt)
    anc = base["ancestors"]
    flags = np.ones(T, bool)
    flags[0] = False
    flags[5] = False  # one step without resampling exercises the carried log-weight tangents
    ref = pf.run_pf(p, y, rng.PRNGKey(6), n, 1.0, "systematic", ancestors=anc, resample_flags=flags, t=tt)
    g = pf.sv_loglik_grad(phi, sigma, beta, y, rng.PRNGKey(6), n, anc, flags, t=tt)

    def ll(ph, sg, bt):
        return pf.run_pf(pf.sv_params(ph, sg, bt), y, rng.PRNGKey(6), n, 1.0, "systematic", ancestors=anc,
                         resample_flags=flags, t=tt)["log_likelihood"]
    h = 1e-5
    fd = [(ll(phi + h, sigma, beta) - ll(phi - h, sigma, beta)) / (2 * h),
          (ll(phi, sigma + h, beta) - ll(phi, sigma - h, beta)) / (2 * h),
          (ll(phi, sigma, beta + h) - ll(phi, sigma, beta - h)) / (2 * h)]
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-5)
    assert abs(ref["log_likelihood"] - ll(phi, sigma, beta)) < 1e-12


def test_resume_from_the_first_state_repeats_the_run():
    """run_particle_filter_for_marginals with initial_sample splits its key straight into the step
    keys; resuming from a run's t = 0 particles with that run's first split key repeats steps 1..T-1
    exactly, and the resumed log-likelihood is the full one minus the t = 0 term."""
    p = _lgssm(5, 3)
    _, y = pf.simulate(p, 9, rng.PRNGKey(4))
    key = rng.PRNGKey(6)
    full = pf.run_pf(p, y, key, 300, 0.6, "systematic")
    rkey = rng.split(key, 2)[0]
    res = pf.run_pf(p, y, rkey, 300, 0.6, "systematic", initial=(full["value"][0], full["log_weight"][0]))
    np.testing.assert_array_equal(res["value"], full["value"])
    np.testing.assert_array_equal(res["ancestors"], full["ancestors"])
    l0 = pf.lse(full["log_weight"][0]) - np.log(300)
    assert abs(res["log_likelihood"] - (full["log_likelihood"] - l0)) < 1e-9


def test_sv_time_steps_scale_the_transition_noise():
    # unit observation times reproduce the unit-step run exactly
    p = pf.sv_params(0.9, 0.4, 0.7)
    y = np.random.default_rng(5).standard_normal((12, 1))
    key = rng.PRNGKey(3)
    a = pf.run_pf(p, y, key, 64, 1.0, "systematic")
    b = pf.run_pf(p, y, key, 64, 1.0, "systematic", t=np.arange(12.0) + 3.0)
    np.testing.assert_array_equal(a["value"], b["value"])
    assert a["log_likelihood"] == b["log_likelihood"]
    # a step of length dt adds variance sigma^2 dt (the transition noise scales with sqrt(dt); a scale of
    # dt, or none, fails): x' = phi x + sigma sqrt(dt) eps from x ~ N(0, 1)
    rs = np.random.default_rng(6)
    x = rs.standard_normal((200000, 1))
    eps = rs.standard_normal((200000, 1))
    for dt in (0.25, 4.0):
        xn = pf.transition(p, x, eps, np.sqrt(dt))
        want = 0.9 ** 2 + 0.4 ** 2 * dt
        assert abs(xn.var() - want) < 4 * want * np.sqrt(2.0 / 200000)


def test_multivariate_sv_reduces_to_1d_and_is_gaussian():
    from scipy.stats import norm
    # dim 1: the multivariate form is the scalar model
    p1 = pf.sv_params(0.9, 0.3, 0.8, m0=0.1, var0=1.5)
    pd = pf.svd_params([0.1], [[1.5]], [[0.9]], [[0.09]], [0.64])
    y = np.random.default_rng(2).standard_normal((9, 1))
    a = pf.run_pf(p1, y, rng.PRNGKey(1), 50, 0.7, "systematic", t=np.arange(9.0) * 0.5)
    b = pf.run_pf(pd, y, rng.PRNGKey(1), 50, 0.7, "systematic", t=np.arange(9.0) * 0.5)
    np.testing.assert_allclose(b["value"], a["value"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b["log_likelihood"], a["log_likelihood"], rtol=1e-12)
    # likelihood: independent normals with standard deviations sqrt(r_k) exp(x_k / 2)
    p3 = pf.svd_params(np.zeros(3), np.eye(3), 0.9 * np.eye(3), np.eye(3), [0.5, 1.0, 2.0])
    x = np.random.default_rng(3).standard_normal((4, 3))
    yv = np.array([0.3, -1.0, 0.7])
    want = -np.sum(norm(0, np.sqrt([0.5, 1.0, 2.0])[None, :] * np.exp(0.5 * x)).logpdf(yv[None, :]), axis=1)
    np.testing.assert_allclose(pf.lik_potential(p3, x, yv), want, rtol=1e-12)
    # transition covariance: A Sx A^T + Q dt (a dropped or misplaced factor fails)
    rs = np.random.default_rng(4)
    A = np.array([[0.9, 0.1], [-0.2, 0.8]])
    Q = np.array([[0.5, 0.2], [0.2, 0.3]])
    p2 = pf.svd_params(np.zeros(2), np.eye(2), A, Q, [1.0, 1.0])
    x = rs.standard_normal((400000, 2))
    e = rs.standard_normal((400000, 2))
    c = np.cov(pf.transition(p2, x, e, np.sqrt(2.5)).T)
    np.testing.assert_allclose(c, A @ A.T + 2.5 * Q, atol=0.02)
