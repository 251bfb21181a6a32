"""GPU logistic regression: fused potential/adjoint vs the FP64 oracle, and the fused HMC iteration
vs the oracle's mocat-faithful iteration from the same state and keys."""
import math

import numpy as np
import pytest
import torch

from oracle import lr as olr
from oracle import mcmc as omcmc
from oracle import rng as orng

pytestmark = pytest.mark.gpu

# measured worst-case relative errors are well inside these; bf16x3 keeps ~16 mantissa bits
TOL = {"fp32": dict(pot=2e-6, grad=2e-5), "bf16x3": dict(pot=2e-5, grad=1e-4), "fp64": dict(pot=2e-7, grad=2e-7)}
# relative error of one potential evaluation (|U| scale) used for the energy difference in alpha
HTOL = {"fp32": 2e-6, "bf16x3": 1e-5, "fp64": 5e-7}


def _case(n, d, C, seed):
    X, y, _ = olr.make_dataset(seed, n, d)
    th = np.random.default_rng(seed + 1).standard_normal((C, d)).astype(np.float32)
    return X, y, th


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp64"])
@pytest.mark.parametrize("n,d,C", [(2000, 50, 64), (77, 7, 33), (5000, 100, 40), (33, 64, 1), (1, 128, 5),
                                   (20000, 100, 70), (1000, 65, 33)])
def test_lr_potential_grad(precision, n, d, C):
    import mocat_amd as m
    X, y, th = _case(n, d, C, n + d)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision=precision)
    pp, lp, gp, gl = [t.cpu().numpy().astype(np.float64) for t in sc.evaluate(torch.tensor(th))]
    U, g, opp, olp = olr.potential_and_grad(th, X, y, 0.1)
    ogl = olr.lik_grad(th, X, y)
    np.testing.assert_allclose(pp, opp, rtol=1e-6)
    scale = np.abs(olr.softplus(th.astype(np.float64) @ X.T.astype(np.float64))).sum(1) + 1.0
    assert np.max(np.abs(lp - olp) / scale) < TOL[precision]["pot"]
    gscale = np.abs(X).sum(0).max() + 1.0
    assert np.max(np.abs(gl - ogl)) / gscale < TOL[precision]["grad"]
    np.testing.assert_allclose(gp, 0.1 * th, rtol=1e-6, atol=1e-7)


def test_lr_adjoint_matches_central_differences():
    import mocat_amd as m
    X, y, th = _case(500, 9, 4, 3)
    sc = m.LogisticRegression(X, y, precision="fp32")
    _, _, _, gl = sc.evaluate(torch.tensor(th))
    gl = gl.cpu().numpy().astype(np.float64)
    eps = 1e-5
    for c in range(4):
        fd = np.array([(olr.lik_potential(th[c] + eps * e, X, y) - olr.lik_potential(th[c] - eps * e, X, y))
                       / (2 * eps) for e in np.eye(9)])
        np.testing.assert_allclose(gl[c], fd, rtol=1e-4, atol=1e-3)


def _gpu_startup(m, precision, n, d, C, seed, leapfrog_steps):
    X, y, _ = olr.make_dataset(seed, n, d)
    sc = m.LogisticRegression(X, y, precision=precision)
    s = m.Underdamped(leapfrog_steps=leapfrog_steps)
    extra = m.cdict(random_key=m.random.PRNGKey(seed))
    state, extra = s.startup(sc, 10, None, extra, n_chains=C, correction=m.RMMetropolis())
    return X, y, sc, s, state, extra


# shapes: the d <= 64 and d > 64 instantiations of every engine, the bench shape (n = 2000, d = 50,
# L = 10), MALA (L = 1) and HMC at d = 100 (the tempered-SMC move), ragged tiles (C not a multiple of 32)
HMC_CASES = [(300, 7, 96, 4), (2000, 50, 64, 10), (500, 65, 40, 1), (1000, 100, 64, 1), (1000, 100, 64, 10),
             (400, 128, 33, 10)]


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp64"])
@pytest.mark.parametrize("n,d,C,L", HMC_CASES)
def test_hmc_iteration_matches_oracle(precision, n, d, C, L):
    import mocat_amd as m
    X, y, sc, s, state, extra = _gpu_startup(m, precision, n, d, C, 5, L)
    # startup parity: prior sample and key tree bit-exact, potentials to tolerance
    ost, okeys = omcmc.startup(orng.PRNGKey(5), C, d, lambda k: np.sqrt(10.0) * orng.normal(k, (d,)),
                               omcmc.lr_pot_grad_fn(X, y, 0.1))
    np.testing.assert_array_equal(extra.random_key.cpu().numpy(), okeys)
    np.testing.assert_allclose(state.value.cpu().numpy(), ost["value"], rtol=2e-6, atol=1e-6)
    # move towards the typical set first (prior draws have potentials in the thousands, where every
    # accept decision is a rounding tie); the comparison then starts from the GPU state
    warm = 40
    s.iterate(sc, state, extra, warm)
    fields = ("value", "momenta", "potential", "grad_potential", "prior_potential", "likelihood_potential",
              "stepsize")
    host = {k: getattr(state, k).cpu().numpy().copy() for k in fields}
    keys0 = extra.random_key.cpu().numpy().copy()
    prev = state.value.cpu().numpy().copy()
    covered = 0
    for it in range(warm + 1, warm + 4):
        s.update(sc, state, extra)
        o, okeys2, a, acc = omcmc.iterate(host, keys0, it, omcmc.lr_pot_grad_fn(X, y, 0.1), L, np.inf, "rm")
        # key tree is integer work: bit-exact
        np.testing.assert_array_equal(extra.random_key.cpu().numpy(), okeys2)
        ga = state.alpha.cpu().numpy()
        gv = state.value.cpu().numpy()
        # alpha = min(1, exp(-dH)): an error e in the energy difference moves alpha by at most
        # alpha * e; e is the engine's potential rounding at this chain's potential scale
        e_h = HTOL[precision] * (4.0 * np.abs(host["potential"]).astype(np.float64) + 10.0 * d)
        bound = 2e-5 + np.maximum(a, ga) * e_h
        assert np.all(np.abs(ga - a) <= bound), np.max(np.abs(ga - a) - bound)
        # accept decisions agree wherever the uniform is not within that bound of alpha
        s2 = orng.split_batch(orng.split_batch(keys0, 2)[:, 0], L + 1)[:, 0]
        u = orng.uniform_batch(orng.split_batch(s2, 2)[:, 1], 1)[:, 0]
        g_acc = ~np.all(gv == prev, axis=1)
        clear = np.abs(u - a) > 2 * bound
        np.testing.assert_array_equal(g_acc[clear], acc[clear])
        # accepted proposals agree to the precision of the gradient engine
        both = clear & acc
        xtol = {"fp32": 2e-4, "bf16x3": 2e-3, "fp64": 1e-4}[precision]
        np.testing.assert_allclose(gv[both], o["value"][both], rtol=xtol, atol=xtol)
        covered += clear.sum()
        # continue both from the GPU state so that rare near-threshold decisions do not compound
        host = {k: getattr(state, k).cpu().numpy().copy() for k in host}
        keys0 = extra.random_key.cpu().numpy().copy()
        prev = gv.copy()
    assert covered >= 0.9 * 3 * C  # the decisions compared cover nearly every chain-iteration


def test_hmc_run_api_and_posterior():
    import mocat_amd as m
    X, y, _ = olr.make_dataset(11, 400, 5)
    res = {}
    for precision in ("fp32", "bf16x3"):
        sc = m.LogisticRegression(X, y, precision=precision)
        chain = m.run(sc, m.Underdamped(leapfrog_steps=5), 400, m.random.PRNGKey(0), n_chains=512,
                      correction=m.RMMetropolis())
        assert chain.value.shape == (400, 512, 5) and chain.potential.shape == (400, 512)
        assert chain.time > 0
        acc = chain.alpha[200:].mean().item()
        assert abs(acc - 0.651) < 0.03, acc
        res[precision] = chain.value[200:].reshape(-1, 5).cpu().numpy()
    # the two kernels sample the same posterior: means agree within Monte Carlo error
    sd = res["fp32"].std(0)
    se = sd * math.sqrt(2.0 / (512 * 200 / 4))  # generous effective-sample-size factor of 4
    assert np.all(np.abs(res["fp32"].mean(0) - res["bf16x3"].mean(0)) < 5 * se + 1e-3)


def test_single_chain_run_matches_mocat_shapes():
    import mocat_amd as m
    X, y, _ = olr.make_dataset(2, 100, 3)
    sc = m.LogisticRegression(X, y)
    chain = m.run(sc, m.Underdamped(leapfrog_steps=2, stepsize=0.1), 50, m.random.PRNGKey(1),
                  correction=m.Metropolis())
    assert chain.value.shape == (50, 3) and chain.potential.shape == (50,)
    assert torch.isfinite(chain.potential).all()
