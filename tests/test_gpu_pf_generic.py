"""The linear-Gaussian model at (d, d_y) outside the fused filters' compiled shapes: the runtime-dimension
hook kernel and the host loop over the library's kernels, against the FP64 oracle on the GPU's own
ancestors (same keys, so the same noise), and the one-step functions against a whole run."""
import numpy as np
import pytest
import torch

from oracle import pf as opf
from oracle import rng as orng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def _model(m, d, dy, seed):
    rs = np.random.default_rng(seed)
    A = 0.8 * np.eye(d) + 0.05 * rs.standard_normal((d, d))
    B = rs.standard_normal((d, d))
    Q = 0.05 * (B @ B.T / d + np.eye(d))
    H = 0.5 * rs.standard_normal((dy, d))
    C = rs.standard_normal((dy, dy))
    R = 0.3 * (C @ C.T / dy + np.eye(dy))
    P0 = np.eye(d) + 0.1 * np.ones((d, d))
    m0 = 0.2 * rs.standard_normal(d)
    model = m.ssm.TimeHomogenousLinearGaussian(m0, P0, A, Q, H, R)
    return model, opf.lgssm_params(m0, P0, A, Q, H, R)


@pytest.mark.parametrize("d,dy,resampling,thr", [(6, 3, "systematic", 0.7), (3, 2, "multinomial", 0.5),
                                                 (12, 7, "categorical", 0.8), (16, 16, "systematic", 1.0)])
def test_generic_shape_same_path_as_fp64_oracle(m, d, dy, resampling, thr):
    model, p = _model(m, d, dy, d * 100 + dy)
    from mocat_amd.ssm import filtering
    assert not filtering._compiled_shape(0, d, dy)
    T, n = 25, 1500
    _, y = opf.simulate(p, T, orng.PRNGKey(d + dy))
    t = np.arange(T, dtype=np.float64)
    key = m.random.PRNGKey(40 + d)
    out = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), torch.tensor(y, dtype=torch.float32),
                                                  t, key, n=n, ess_threshold=thr, resampling=resampling)
    assert out.value.shape == (T, n, d) and out.log_weight.shape == (T, n) and out.ancestors.shape == (T, n)
    ess = out.ess.cpu().numpy()
    flags = np.zeros(T, bool)
    flags[1:] = ess[:-1] < thr * n
    assert flags.any()
    anc = out.ancestors.cpu().numpy()
    o = opf.run_pf(p, y, key.cpu().numpy(), n, thr, resampling, ancestors=anc, resample_flags=flags)
    ll = out.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) <= 1e-4 * (1 + abs(o["log_likelihood"])), (ll, o["log_likelihood"])
    np.testing.assert_allclose(out.value[0].cpu().numpy(), o["value"][0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(out.value.cpu().numpy(), o["value"], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(ess, o["ess"], rtol=1e-3)
    for k in range(1, T):
        if not flags[k]:
            assert np.array_equal(anc[k], np.arange(n))
    # systematic / multinomial steps: the oracle's resampler, given the GPU's log-weights and the step
    # key, draws the same ancestors (up to the weights' f32 rounding at interval ends)
    if resampling != "categorical":
        lw = out.log_weight.cpu().numpy().astype(np.float64)
        ikeys = orng.split(orng.split(key.cpu().numpy(), 2)[0], T - 1)
        agree = [(opf.resample_indices(np.exp(lw[k - 1] - lw[k - 1].max()).astype(np.float32), ikeys[k - 1],
                                       resampling) == anc[k]).mean() for k in range(1, T) if flags[k]]
        assert np.mean(agree) > 0.995, agree


def test_generic_shape_hooks_and_batches(m):
    d, dy = 6, 3
    model, p = _model(m, d, dy, 7)
    ks = m.random.split(m.random.PRNGKey(2), 64)
    eps = np.stack([orng.normal(k, (d,)) for k in ks.cpu().numpy()]).astype(np.float64)
    x0 = model.initial_sample(0.0, ks)
    np.testing.assert_allclose(x0.cpu().numpy(), opf.init_state(p, eps), rtol=1e-5, atol=1e-5)
    x1 = model.transition_sample(x0, 0.0, 1.0, ks)
    np.testing.assert_allclose(x1.cpu().numpy(), opf.transition(p, x0.cpu().numpy().astype(np.float64), eps),
                               rtol=1e-5, atol=1e-5)
    yv = np.linspace(-1, 1, dy)
    lik = model.likelihood_potential(x1, torch.tensor(yv, dtype=torch.float32, device="cuda"), 1.0)
    np.testing.assert_allclose(lik.cpu().numpy(), opf.lik_potential(p, x1.cpu().numpy().astype(np.float64), yv),
                               rtol=1e-5, atol=1e-4)
    # batched filters of this shape: each as its own run
    _, y = opf.simulate(p, 12, orng.PRNGKey(1))
    keys = [k for k in m.random.split(m.random.PRNGKey(8), 3)]
    ns = [200, 1300, 5000]
    res = m.ssm.run_particle_filters([model] * 3, [y] * 3, keys, ns, 0.6, "systematic", store=True)
    for r, k, n in zip(res, keys, ns):
        one = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), torch.tensor(y), np.arange(12.0),
                                                      k, n=n, ess_threshold=0.6, resampling="systematic")
        assert r.log_likelihood.item() == one.log_likelihood.item()
        assert torch.equal(r.value, one.value) and r.final_value.shape == (n, d)


@pytest.mark.parametrize("d,dy,resampling", [(6, 3, "systematic"), (16, 9, "multinomial")])
def test_generic_shape_optimal_filter_same_path_as_fp64_oracle(m, d, dy, resampling):
    """OptimalNonLinearGaussianParticleFilter at a runtime-dimension shape vs the FP64 oracle on the
    GPU's own resampling decisions and ancestors (zero initial weights; log p(y_1:T | y_0))."""
    from oracle import nlg as onlg
    model, p = _model(m, d, dy, 3 * d + dy)
    q = onlg.startup(model.initial_mean, model.initial_covariance, model.transition_matrix,
                     model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)
    T, n, thr = 20, 1200, 0.7
    _, y = opf.simulate(p, T, orng.PRNGKey(d))
    key = m.random.PRNGKey(9)
    out = m.ssm.run_particle_filter_for_marginals(model, m.ssm.OptimalNonLinearGaussianParticleFilter(),
                                                  torch.tensor(y, dtype=torch.float32), np.arange(float(T)), key,
                                                  n=n, ess_threshold=thr, resampling=resampling)
    ess = out.ess.cpu().numpy()
    flags = np.zeros(T, bool)
    flags[1:] = ess[:-1] < thr * n
    o = onlg.run_optimal_pf(q, y, key.cpu().numpy(), n, thr, resampling, ancestors=out.ancestors.cpu().numpy(),
                            resample_flags=flags)
    np.testing.assert_allclose(out.value[0].cpu().numpy(), o["value"][0], rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(out.log_weight[0].cpu().numpy(), np.zeros(n, np.float32))
    np.testing.assert_allclose(out.value[-1].cpu().numpy(), o["value"][-1], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(out.log_weight[-1].cpu().numpy(), o["log_weight"][-1], rtol=1e-4, atol=1e-3)
    ll = out.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]) + 1e-4, (ll, o["log_likelihood"])
    np.testing.assert_allclose(ess, o["ess"], rtol=1e-3)


def test_generic_shape_one_step_functions_repeat_the_run(m):
    d, dy, T, n = 7, 2, 8, 900
    model, p = _model(m, d, dy, 11)
    _, yn = opf.simulate(p, T, orng.PRNGKey(5))
    y = torch.tensor(yn, dtype=torch.float32)
    t = np.arange(T, dtype=np.float64)
    pf = m.ssm.BootstrapFilter()
    key = m.random.PRNGKey(13)
    full = m.ssm.run_particle_filter_for_marginals(model, pf, y, t, key, n=n, ess_threshold=0.6,
                                                   resampling="categorical")
    rkey, sub = m.random.split(key, 2)
    q = m.ssm.initiate_particles(model, pf, n, sub, y[0], float(t[0]))
    assert torch.equal(q.value[0], full.value[0]) and torch.equal(q.log_weight[0], full.log_weight[0])
    int_keys = m.random.split(rkey, T - 1)
    for k in range(1, T):
        q = m.ssm.propagate_particle_filter(model, pf, q, y[k], float(t[k]), int_keys[k - 1], ess_threshold=0.6,
                                            resample_full=False)
        assert torch.equal(q.value[k], full.value[k]) and torch.equal(q.log_weight[k], full.log_weight[k])
    # resumed from the run's t = 0 sample: the rest of the run, and the log-likelihood without the t = 0 term
    res = m.ssm.run_particle_filter_for_marginals(model, pf, y[1:], t[1:], rkey,
                                                  initial_sample=m.ssm.initiate_particles(model, pf, n, sub, y[0],
                                                                                          float(t[0])),
                                                  ess_threshold=0.6, resampling="categorical")
    assert torch.equal(res.value, full.value)
    lw0 = full.log_weight[0].double()
    l0 = (torch.logsumexp(lw0, 0) - np.log(n)).item()
    assert abs(res.log_likelihood.item() - (full.log_likelihood.item() - l0)) < 1e-9 * (1 + abs(l0))


@pytest.mark.parametrize("d,dy", [(6, 3), (11, 4)])
def test_generic_shape_backward_simulation_vs_oracle(m, d, dy):
    """Backward simulation at a d without a compiled kernel: full resampling and rejection rounds, every
    step's indices equal to the oracle's on the same keys (given the GPU's trajectories at the next
    time), except draws whose best two scores (or acceptance tests) lie within 1e-4."""
    from oracle import ffbsi as of
    model, p = _model(m, d, dy, 5 * d)
    T, n, n_samps = 6, 400, 96
    _, y = opf.simulate(p, T, orng.PRNGKey(3))
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), torch.tensor(y, dtype=torch.float32),
                                                   np.arange(float(T)), m.random.PRNGKey(4), n=n, ess_threshold=0.5)
    key = m.random.PRNGKey(5)
    res = m.ssm.backward_simulation(model, marg, key, n_samps)
    assert res.value.shape == (T, n_samps, d)
    pot = of.lgssm_transition(model.transition_matrix, model.transition_covariance)
    vals = marg.value.cpu().numpy().astype(np.float64)
    lw = marg.log_weight.cpu().numpy().astype(np.float64)
    out, anc = res.value.cpu().numpy(), res.ancestors.cpu().numpy()
    tk = orng.split(key.cpu().numpy(), T)
    ties = []
    for t in range(T - 2, -1, -1):
        want, margin = of.backward_step_margin(pot, vals[t], lw[t], out[t + 1], tk[t])
        decided = margin > 1e-4
        np.testing.assert_array_equal(anc[t][decided], want[decided])
        ties.append((~decided).mean())
    assert max(ties) <= 0.02, ties
    np.testing.assert_array_equal(out, np.take_along_axis(vals, anc[:, :, None].astype(np.int64), 1).astype(np.float32))
    # rejection rounds (backward.py:42-194) at this d: the same proposals, acceptances, bound raises and
    # final full draws as the oracle on the same keys, except near-ties; the evaluation counts follow
    for max_rej, bound in ((1, 0.0), (12, 0.0)):
        res = m.ssm.backward_simulation(model, marg, key, n_samps, maximum_rejections=max_rej,
                                        transition_dens_bound_parameter=bound, bound_inflation=1.01)
        out, anc = res.value.cpu().numpy(), res.ancestors.cpu().numpy()
        ties = []
        for t in range(T - 2, -1, -1):
            idx, evals, tie = of.rejection_step(pot, vals[t], lw[t], out[t + 1], tk[t], max_rej, bound, 1.01, 1e-4)
            np.testing.assert_array_equal(anc[t][~tie], idx[~tie])
            ties.append(tie.mean())
            got = res.num_transition_evals[t + 1].item()
            assert abs(got - evals) <= 0.02 * evals + n, (got, evals)
        assert max(ties) <= 0.02, ties
        np.testing.assert_array_equal(out, np.take_along_axis(vals, anc[:, :, None].astype(np.int64), 1)
                                      .astype(np.float32))


@pytest.mark.parametrize("d,dy,n", [(6, 3, 3000), (13, 5, 2)])
def test_generic_shape_enkf_same_path_as_fp64_oracle(m, d, dy, n):
    """EnsembleKalmanFilter at a runtime-dimension shape: the same keys and noise as the FP64 oracle;
    ensembles and gains at every step (n = 2 is the smallest ensemble with a covariance)."""
    from oracle import nlg as onlg
    model, p = _model(m, d, dy, 7 * d + dy)
    q = onlg.startup(model.initial_mean, model.initial_covariance, model.transition_matrix,
                     model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)
    T = 8
    _, y = opf.simulate(p, T, orng.PRNGKey(d))
    key = m.random.PRNGKey(17)
    out = m.ssm.run_particle_filter_for_marginals(model, m.ssm.EnsembleKalmanFilter(), torch.tensor(y, dtype=torch.float32),
                                                  np.arange(float(T)), key, n=n)
    o = onlg.run_enkf(q, y, key.cpu().numpy(), n)
    got = out.value.cpu().numpy()
    assert got.shape == (T, n, d)
    np.testing.assert_allclose(got[0], o["value"][0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got, o["value"], rtol=1e-4, atol=2e-4)
    np.testing.assert_allclose(out.gain.cpu().numpy(), o["gain"], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(out.final_value.cpu().numpy(), o["value"][-1], rtol=1e-4, atol=2e-4)
