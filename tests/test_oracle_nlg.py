"""Pins the optimal-proposal particle filter and ensemble Kalman filter oracles."""
import numpy as np
from scipy.stats import multivariate_normal

from oracle import kalman, nlg, pf, rng


def _spd(rs, d, scale):
    B = rs.normal(0, 1, (d, d))
    return scale * (B @ B.T / d + 0.5 * np.eye(d))


def _model(seed, d=3, dy=2, diag=True):
    rs = np.random.default_rng(seed)
    A = 0.8 * np.eye(d) + rs.normal(0, 0.15, (d, d))
    Q = _spd(rs, d, 0.3)
    H = rs.normal(0, 0.7, (dy, d))
    P0 = np.diag(rs.uniform(0.5, 2.0, d)) if diag else _spd(rs, d, 1.0)
    R = np.diag(rs.uniform(0.2, 1.0, dy)) if diag else _spd(rs, dy, 0.5)
    return dict(m0=rs.normal(0, 1, d), P0=P0, A=A, Q=Q, H=H, R=R)


def _startup(m):
    return nlg.startup(m["m0"], m["P0"], m["A"], m["Q"], m["H"], m["R"])


def test_optimal_weight_times_proposal_is_transition_times_likelihood():
    # q(x' | x, y) exp(w(x, y)) = f(x' | x) g(y | x') for every x': pins the weight precision and its
    # normaliser (full R and Q: the proposal and weight use R only through LR LR^T)
    m = _model(1, diag=False)
    p = _startup(m)
    rs = np.random.default_rng(2)
    xp = rs.normal(0, 1, (6, 3))
    y = rs.normal(0, 1, 2)
    z = rs.normal(0, 1, (6, 3))
    x, w = nlg.optimal_step(p, xp, y, z)
    Qi, Ri = np.linalg.inv(m["Q"]), np.linalg.inv(m["R"])
    cov_q = np.linalg.inv(Qi + m["H"].T @ Ri @ m["H"])  # information form, not the gain form of mocat
    for i in range(6):
        mean_q = cov_q @ (Qi @ m["A"] @ xp[i] + m["H"].T @ Ri @ y)
        lq = multivariate_normal(mean_q, cov_q).logpdf(x[i])
        lf = multivariate_normal(m["A"] @ xp[i], m["Q"]).logpdf(x[i])
        lg = multivariate_normal(m["H"] @ x[i], m["R"]).logpdf(y)
        assert abs(lq + w[i] - lf - lg) < 1e-9


def test_optimal_proposal_is_the_conditional_gaussian():
    m = _model(3, diag=False)
    p = _startup(m)
    xp = np.array([[0.3, -1.0, 2.0]])
    y = np.array([0.5, -0.2])
    Qi, Ri = np.linalg.inv(m["Q"]), np.linalg.inv(m["R"])
    cov_q = np.linalg.inv(Qi + m["H"].T @ Ri @ m["H"])
    mean_q = cov_q @ (Qi @ m["A"] @ xp[0] + m["H"].T @ Ri @ y)
    x0, _ = nlg.optimal_step(p, xp, y, np.zeros((1, 3)))
    np.testing.assert_allclose(x0[0], mean_q, atol=1e-12)
    # unit noise vectors give the columns of the proposal covariance square root
    cols = np.stack([nlg.optimal_step(p, xp, y, np.eye(3)[k][None, :])[0][0] - x0[0] for k in range(3)], 1)
    np.testing.assert_allclose(cols @ cols.T, cov_q, atol=1e-12)


def test_initial_sample_is_the_posterior_of_x0_given_y0():
    # diagonal P0 and R (where mocat's initial conditioned precision is the posterior precision)
    m = _model(4, diag=True)
    p = _startup(m)
    y0 = np.array([1.0, -0.5])
    Pi, Ri = np.linalg.inv(m["P0"]), np.linalg.inv(m["R"])
    cov0 = np.linalg.inv(Pi + m["H"].T @ Ri @ m["H"])
    mean0 = cov0 @ (Pi @ m["m0"] + m["H"].T @ Ri @ y0)
    x0 = nlg.initial(p, y0, np.zeros((1, 3)))[0]
    np.testing.assert_allclose(x0, mean0, atol=1e-12)
    rows = nlg.initial(p, y0, np.eye(3)) - x0[None, :]
    np.testing.assert_allclose(rows.T @ rows, cov0, atol=1e-12)


def _kalman(m, y):
    return kalman.loglik(m["m0"], m["P0"], m["A"], m["Q"], m["H"], m["R"], y)


def _simulate(m, T, key):
    q = pf.lgssm_params(m["m0"], m["P0"], m["A"], m["Q"], m["H"], m["R"])
    return pf.simulate(q, T, key)[1]


def test_optimal_pf_converges_to_kalman():
    m = _model(5, d=2, dy=2, diag=True)
    p = _startup(m)
    y = _simulate(m, 8, rng.PRNGKey(1))
    n = 20000
    out = nlg.run_optimal_pf(p, y, rng.PRNGKey(2), n, resampling="systematic")
    ll, means, covs = _kalman(m, y)
    ll0 = multivariate_normal(m["H"] @ m["m0"], m["H"] @ m["P0"] @ m["H"].T + m["R"]).logpdf(y[0])
    for t in range(8):
        w = np.exp(out["log_weight"][t] - pf.lse(out["log_weight"][t]))
        mean = w @ out["value"][t]
        sd = np.sqrt(np.diag(covs[t]) / out["ess"][t])
        assert np.all(np.abs(mean - means[t]) < 5 * sd), (t, mean, means[t])
    # zero initial weights: the estimate is of log p(y_1:T | y_0)
    assert abs(out["log_likelihood"] - (ll - ll0)) < 0.05
    assert abs(out["ess"][0] - n) < 1e-9 * n


def test_enkf_converges_to_kalman():
    m = _model(6, d=3, dy=2, diag=True)
    p = _startup(m)
    y = _simulate(m, 6, rng.PRNGKey(3))
    n = 40000
    out = nlg.run_enkf(p, y, rng.PRNGKey(4), n)
    _, means, covs = _kalman(m, y)
    for t in range(6):
        x = out["value"][t]
        sd = np.sqrt(np.diag(covs[t]) / n)
        assert np.all(np.abs(x.mean(0) - means[t]) < 6 * sd), (t, x.mean(0), means[t])
        np.testing.assert_allclose(np.cov(x.T), covs[t], atol=0.03 * np.abs(covs[t]).max())


def test_enkf_step_closed_forms():
    rs = np.random.default_rng(7)
    # 1-D: K = P h / (h^2 P + r) with P the unbiased sample variance
    a, q, h, r = 0.9, 0.4, 1.3, 0.25
    p = nlg.startup(np.zeros(1), np.eye(1), [[a]], [[q]], [[h]], [[r]])
    xp, z1, z2 = rs.normal(0, 1, (50, 1)), rs.normal(0, 1, (50, 1)), rs.normal(0, 1, (50, 1))
    x, K = nlg.enkf_step(p, xp, np.array([0.7]), z1, z2)
    mx = a * xp[:, 0] + np.sqrt(q) * z1[:, 0]
    P = np.var(mx, ddof=1)
    k = P * h / (h * h * P + r)
    np.testing.assert_allclose(K[0, 0], k, rtol=1e-12)
    np.testing.assert_allclose(x[:, 0], mx + k * (0.7 - h * mx - np.sqrt(r) * z2[:, 0]), rtol=1e-12, atol=1e-12)
    # d = 3, dy = 2: K = P H^T (H P H^T + R)^-1 with P = np.cov of the forecast ensemble
    m = _model(8, diag=False)
    m["P0"] = np.eye(3)  # the step does not read P0
    p = _startup(m)
    xp, z1, z2 = rs.normal(0, 1, (40, 3)), rs.normal(0, 1, (40, 3)), rs.normal(0, 1, (40, 2))
    x, K = nlg.enkf_step(p, xp, np.array([0.1, 0.2]), z1, z2)
    mx = xp @ m["A"].T + z1 @ np.linalg.cholesky(m["Q"]).T
    P = np.cov(mx.T)
    np.testing.assert_allclose(K, P @ m["H"].T @ np.linalg.inv(m["H"] @ P @ m["H"].T + m["R"]), rtol=1e-10)


def test_shared_key_draws():
    # every particle's noise comes from the first key of split(key, n), flattened row-major
    key = rng.PRNGKey(9)
    z = nlg.shared_normals(key, 5, 3)
    k0 = rng.split(key, 5)[0]
    np.testing.assert_array_equal(z.ravel(), rng.normal(k0, (15,)).astype(np.float64))
    # and the draw depends on n (split(key, n)[0] changes with n)
    assert not np.array_equal(nlg.shared_normals(key, 6, 3)[:5], z)
