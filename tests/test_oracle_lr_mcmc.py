"""Pins the LR and MCMC oracles: finite differences, the Bernoulli log-pmf, mocat's leapfrog KAT."""
import numpy as np

from oracle import lr, mcmc, rng


def test_lr_grad_central_differences():
    X, y, _ = lr.make_dataset(0, 300, 7)
    th = np.random.default_rng(1).standard_normal(7)
    g = lr.lik_grad(th, X, y)
    eps = 1e-6
    fd = np.array([(lr.lik_potential(th + eps * e, X, y) - lr.lik_potential(th - eps * e, X, y)) / (2 * eps)
                   for e in np.eye(7)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_lr_potential_is_negative_bernoulli_loglik():
    X, y, _ = lr.make_dataset(3, 200, 4)
    th = np.random.default_rng(2).standard_normal(4)
    p = 1 / (1 + np.exp(-(X.astype(np.float64) @ th)))
    ll = np.sum(y * np.log(p) + (1 - y) * np.log1p(-p))
    np.testing.assert_allclose(lr.lik_potential(th, X, y), -ll, rtol=1e-10)
    # prior = -log N(theta; 0, I/prec)
    prec = 0.1
    lp = -0.5 * prec * th @ th - 0.5 * 4 * np.log(2 * np.pi / prec)
    np.testing.assert_allclose(lr.prior_potential(th, prec), -lp, rtol=1e-12)


def test_leapfrog_reference_kat():
    # mocat/src/tests/test_utils.py:121-147: U = 0.5|x|^2 (lik grad = x), stepsize 0.1, 3 steps
    def grad_fn(x):
        return np.float32(0.5) * np.sum(x * x), x.astype(np.float32)
    x, p, g, _ = mcmc.leapfrog(grad_fn, np.zeros(2), np.ones(2), np.array([1.0, 2.0]), 0.1, 3)
    np.testing.assert_array_equal(x, np.array([0.28120947, 0.266409], np.float32))
    np.testing.assert_array_equal(g, np.array([0.28120947, 0.266409], np.float32))
    np.testing.assert_array_equal(p, np.array([0.9075344, 0.8597696], np.float32))


def test_hmc_gaussian_moments():
    # HMC on N(mu, diag(s2)): posterior moments within Monte Carlo error
    mu = np.array([1.0, -2.0], np.float32)
    s2 = np.array([0.5, 2.0], np.float32)

    def pg(x):
        u = 0.5 * np.sum((x - mu) ** 2 / s2, axis=-1)
        return u, (x - mu) / s2, np.zeros_like(u), u

    # fixed stepsize (Metropolis): exact target, moments within MC error
    st, keys = mcmc.startup(rng.PRNGKey(0), 256, 2, lambda k: rng.normal(k, (2,)), pg, stepsize=0.5)
    xs = []
    for it in range(1, 301):
        st, keys, a, acc = mcmc.iterate(st, keys, it, pg, 5, np.inf, "mh")
        if it > 100:
            xs.append(st["value"].copy())
    xs = np.stack(xs)
    np.testing.assert_allclose(xs.mean((0, 1)), mu, atol=0.03)
    np.testing.assert_allclose(xs.var((0, 1)), s2, rtol=0.04)
    # Robbins-Monro: mean acceptance driven to the 0.651 target
    st, keys = mcmc.startup(rng.PRNGKey(1), 256, 2, lambda k: rng.normal(k, (2,)), pg)
    acc_hist = []
    for it in range(1, 301):
        st, keys, a, acc = mcmc.iterate(st, keys, it, pg, 5, np.inf, "rm", 0.651)
        acc_hist.append(a.mean())
    assert abs(np.mean(acc_hist[150:]) - 0.651) < 0.02
