"""GPU particle filters and resampling vs the oracle (same keys => same random path)."""
import numpy as np
import pytest
import torch

from oracle import kalman as okal
from oracle import pf as opf
from oracle import rng as orng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


@pytest.mark.parametrize("method", ["systematic", "multinomial"])
@pytest.mark.parametrize("n", [1, 2, 7, 1000, 4096, 65539])
def test_resampling_bit_exact(m, method, n):
    rs = np.random.default_rng(n)
    w = (rs.uniform(0, 1, n) ** 3 * (rs.uniform(size=n) > 0.2)).astype(np.float32)
    w[rs.integers(0, n)] = 1.0
    key = orng.PRNGKey(n + 1)
    got = m.ssm.resample_indices(torch.tensor(w, device="cuda"), m.random.PRNGKey(n + 1), method).cpu().numpy()
    want = opf.resample_indices(w, key, method)
    np.testing.assert_array_equal(got, want)


def test_resampling_edge_cases(m):
    # all weight on one particle; tiny and zero weights; all-zero weights select the last index
    for w in (np.array([0, 0, 5, 0], np.float32), np.array([1e-30, 1.0, 1e-38, 0], np.float32),
              np.zeros(5, np.float32)):
        for method in ("systematic", "multinomial"):
            got = m.ssm.resample_indices(torch.tensor(w, device="cuda"), m.random.PRNGKey(9), method).cpu().numpy()
            np.testing.assert_array_equal(got, opf.resample_indices(w, orng.PRNGKey(9), method))


# Gumbel scores -log(-log u) + lw are formed in single precision on the GPU (a few ulp of |score| <= ~25)
# and in FP64 by the oracle: a draw whose best two scores are closer than TIE can have another first
# maximum in f32. Every other draw must agree exactly, and such near-ties must stay rare.
TIE = 2e-5


@pytest.mark.parametrize("n,n_out", [(64, 4096), (4096, 4096), (1000, 4099)])
def test_categorical_matches_gumbel_max(m, n, n_out):
    lw = np.log(np.random.default_rng(n).uniform(0.01, 1, n)).astype(np.float32)
    got = m.ssm.categorical(m.random.PRNGKey(3), torch.tensor(lw, device="cuda"), n_out).cpu().numpy()
    want, margin = orng.categorical_margin(orng.PRNGKey(3), lw, n_out)
    decided = margin > TIE
    np.testing.assert_array_equal(got[decided], want[decided])
    assert (~decided).sum() <= max(2, n_out // 1000), (~decided).sum()


def _lgssm_np(model):
    return opf.lgssm_params(model.initial_mean, model.initial_covariance, model.transition_matrix,
                            model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)


@pytest.mark.parametrize("resampling,thr", [("systematic", 1.0), ("multinomial", 0.5), ("systematic", 0.0)])
def test_lgssm_filter_same_path_as_fp64_oracle(m, resampling, thr):
    """GPU f32 filter vs the FP64 restatement on the same random path: same keys, and the GPU's own
    resampling decisions and ancestors replayed (resampling is chaotic in the weights' last bits, so the
    path is conditioned on it); the resampling itself is checked step by step against the oracle."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, 30)
    n = 1024
    out = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), y, t, m.random.PRNGKey(1),
                                                  n=n, ess_threshold=thr, resampling=resampling)
    assert out.value.shape == (30, n, 4) and out.log_weight.shape == (30, n) and out.ess.shape == (30,)
    anc = out.ancestors.cpu().numpy()
    ess = out.ess.cpu().numpy()
    flags = np.zeros(30, bool)
    flags[1:] = ess[:-1] < thr * n
    p = _lgssm_np(model)
    o = opf.run_pf(p, y.cpu().numpy(), orng.PRNGKey(1), n, thr, resampling, ancestors=anc, resample_flags=flags)
    np.testing.assert_allclose(out.value[0].cpu().numpy(), o["value"][0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(out.value[-1].cpu().numpy(), o["value"][-1], rtol=1e-3, atol=1e-3)
    ll = out.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]), (ll, o["log_likelihood"])
    np.testing.assert_allclose(ess, o["ess"], rtol=1e-3)
    # each resampling step: the oracle, given the GPU's log-weights and the step key, draws the same ancestors
    lw = out.log_weight.cpu().numpy().astype(np.float64)
    rkey = orng.split(orng.PRNGKey(1), 2)[0]
    ikeys = orng.split(rkey, 29)
    agree = []
    for k in range(1, 30):
        if flags[k]:
            w = np.exp(lw[k - 1] - opf.lse(lw[k - 1])).astype(np.float32)
            agree.append((opf.resample_indices(w, ikeys[k - 1], resampling) == anc[k]).mean())
        else:
            assert np.array_equal(anc[k], np.arange(n))
    if agree:
        assert np.mean(agree) > 0.995, agree


def test_config1_loglik_vs_kalman(m):
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, 100)
    R = 32
    keys = m.random.split(m.random.PRNGKey(7), R)
    res = m.ssm.run_particle_filters([model] * R, [y.cpu().numpy()] * R, [k for k in keys], [4096] * R,
                                     ess_threshold=1.0, resampling="systematic")
    lls = np.array([r.log_likelihood.item() for r in res])
    p = _lgssm_np(model)
    kll, _, _ = okal.loglik(p["m0"], np.eye(4), p["A"], p["LQ"] @ p["LQ"].T, p["H"],
                            np.linalg.inv(p["SR"] @ p["SR"].T), y.cpu().numpy().astype(np.float64))
    gk = m.ssm.run_kalman_filter_for_marginals(model, y).log_likelihood.item()
    assert abs(gk - kll) < 1e-8 * abs(kll)
    mean, se = lls.mean(), lls.std(ddof=1) / np.sqrt(R)
    # bootstrap-filter log-likelihoods are biased low by ~Var/2; allow 4 SE plus that bias
    assert abs(mean - kll) < 4 * se + 0.5 * lls.var(), (mean, se, kll)


def test_sv_batched_uneven_same_path(m):
    from mocat_amd import datasets
    models, ys, ns = datasets.sv_batch_config(3, n_models=6, n_min=64, n_max=3000, t_min=5, t_max=60)
    keys = m.random.split(m.random.PRNGKey(11), 6)
    res = m.ssm.run_particle_filters(models, ys, [k for k in keys], ns, ess_threshold=1.0,
                                     resampling="systematic", store=True)
    for i in range(6):
        p = opf.sv_params(models[i].phi, models[i].sigma, models[i].beta)
        T = len(ys[i])
        flags = np.ones(T, bool)
        flags[0] = False
        o = opf.run_pf(p, ys[i], keys[i].cpu().numpy(), ns[i], 1.0, "systematic",
                       ancestors=res[i].ancestors.cpu().numpy(), resample_flags=flags)
        ll = res[i].log_likelihood.item()
        assert abs(ll - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]), (i, ll, o["log_likelihood"])
        assert res[i].value.shape == (T, ns[i], 1)


def test_categorical_resampling_filter_runs(m):
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, 10)
    out = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), y, t, m.random.PRNGKey(2),
                                                  n=256, ess_threshold=0.5)
    assert torch.isfinite(out.log_likelihood)
    o = opf.run_pf(_lgssm_np(model), y.cpu().numpy(), orng.PRNGKey(2), 256, 0.5, "systematic")
    # different resampler, same model: log-likelihoods agree to Monte Carlo error
    assert abs(out.log_likelihood.item() - o["log_likelihood"]) < 5.0


def test_sv_loglik_gradient_fixed_index(m):
    """d log Z / d(phi, sigma, beta) of the fixed-resampling-index estimator: GPU forward tangents vs the
    FP64 oracle on the GPU's own path (its ancestors replayed); the oracle is pinned by finite
    differences."""
    from mocat_amd import datasets
    models, ys, ns = datasets.sv_batch_config(7, n_models=5, n_min=128, n_max=2000, t_min=10, t_max=80)
    keys = m.random.split(m.random.PRNGKey(21), 5)
    res = m.ssm.run_particle_filters(models, ys, [k for k in keys], ns, ess_threshold=0.7,
                                     resampling="systematic", store=True, grad=True)
    plain = m.ssm.run_particle_filters(models, ys, [k for k in keys], ns, ess_threshold=0.7,
                                       resampling="systematic")
    for i in range(5):
        T = len(ys[i])
        ess = res[i].ess.cpu().numpy()
        flags = np.zeros(T, bool)
        flags[1:] = ess[:-1] < 0.7 * ns[i]
        g = res[i].log_likelihood_grad.cpu().numpy()
        og = opf.sv_loglik_grad(models[i].phi, models[i].sigma, models[i].beta, ys[i][:, 0], keys[i].cpu().numpy(),
                                ns[i], res[i].ancestors.cpu().numpy(), flags)
        np.testing.assert_allclose(g, og, rtol=2e-3, atol=2e-3 * (1 + np.abs(og).max()))
        # the gradient path leaves the log-likelihood untouched
        assert res[i].log_likelihood.item() == plain[i].log_likelihood.item()


@pytest.mark.parametrize("large_n", [1 << 40, 1])
def test_sv_irregular_times_vs_fp64_oracle(m, large_n):
    """Observation times with uneven gaps: the stochastic-volatility transition noise scales with
    sqrt(dt) (mocat/src/ssm/scenarios/stochastic_volatility.py:86-93). Both filter paths vs the FP64
    oracle on the GPU's own ancestors (log-likelihood, particles, and the gradient's forward tangents)."""
    from mocat_amd import datasets
    models, ys, _ = datasets.sv_batch_config(12, n_models=3, n_min=300, n_max=3000, t_min=20, t_max=60)
    ns = [300, 1025, 3000]
    rs = np.random.default_rng(3)
    ts = [np.cumsum(rs.uniform(0.1, 4.0, len(y))) for y in ys]
    keys = [k for k in m.random.split(m.random.PRNGKey(8), 3)]
    res = m.ssm.run_particle_filters(models, ys, keys, ns, ess_threshold=0.7, resampling="systematic", store=True,
                                     grad=True, large_n=large_n, ts=ts)
    unit = m.ssm.run_particle_filters(models, ys, keys, ns, ess_threshold=0.7, resampling="systematic",
                                      large_n=large_n)
    for i in range(3):
        T = len(ys[i])
        ess = res[i].ess.cpu().numpy()
        flags = np.zeros(T, bool)
        flags[1:] = ess[:-1] < 0.7 * ns[i]
        p = opf.sv_params(models[i].phi, models[i].sigma, models[i].beta)
        anc = res[i].ancestors.cpu().numpy()
        o = opf.run_pf(p, ys[i], keys[i].cpu().numpy(), ns[i], 0.7, "systematic", ancestors=anc,
                       resample_flags=flags, t=ts[i])
        ll = res[i].log_likelihood.item()
        assert abs(ll - o["log_likelihood"]) < 1e-4 * (1 + abs(ll)), (i, ll, o["log_likelihood"])
        np.testing.assert_allclose(res[i].value.cpu().numpy(), o["value"], rtol=1e-4, atol=1e-4)
        og = opf.sv_loglik_grad(models[i].phi, models[i].sigma, models[i].beta, ys[i][:, 0], keys[i].cpu().numpy(),
                                ns[i], anc, flags, t=ts[i])
        g = res[i].log_likelihood_grad.cpu().numpy()
        np.testing.assert_allclose(g, og, rtol=2e-3, atol=2e-3 * (1 + np.abs(og).max()))
        assert ll != unit[i].log_likelihood.item()  # the times change the run


def _svd_case(d, seed):
    rs = np.random.default_rng(seed)
    A = 0.85 * np.eye(d) + 0.05 * rs.standard_normal((d, d))
    B = rs.standard_normal((d, d))
    Q = 0.1 * (B @ B.T / d + 0.5 * np.eye(d))
    P0 = np.eye(d) + 0.2 * np.ones((d, d))
    m0 = 0.1 * rs.standard_normal(d)
    r = rs.uniform(0.4, 1.5, d)
    return m0, P0, A, Q, r


@pytest.mark.parametrize("large_n", [1 << 40, 1])
@pytest.mark.parametrize("d", [2, 3, 4])
def test_multivariate_sv_vs_fp64_oracle(m, d, large_n):
    """mocat's StochasticVolatility with dim > 1 (full transition matrix and covariance, diagonal
    likelihood covariance) and irregular times, both filter paths, vs the FP64 oracle on the GPU's
    own ancestors; the model's hooks draw what the oracle's model functions give for the same keys."""
    m0, P0, A, Q, r = _svd_case(d, d)
    model = m.ssm.StochasticVolatility(d, m0, P0, A, Q, r)
    p = opf.svd_params(m0, P0, A, Q, r)
    T, n = 40, 1500
    _, y = opf.simulate(p, T, orng.PRNGKey(d))
    t = np.cumsum(np.random.default_rng(d).uniform(0.2, 2.0, T))
    key = m.random.PRNGKey(30 + d)
    res = m.ssm.run_particle_filters([model], [y.astype(np.float32)], [key], [n], ess_threshold=0.7,
                                     resampling="systematic", store=True, large_n=large_n, ts=[t])[0]
    ess = res.ess.cpu().numpy()
    flags = np.zeros(T, bool)
    flags[1:] = ess[:-1] < 0.7 * n
    o = opf.run_pf(p, y, key.cpu().numpy(), n, 0.7, "systematic", ancestors=res.ancestors.cpu().numpy(),
                   resample_flags=flags, t=t)
    ll = res.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) < 1e-4 * (1 + abs(ll)), (ll, o["log_likelihood"])
    np.testing.assert_allclose(res.value.cpu().numpy(), o["value"], rtol=1e-4, atol=1e-4)
    # hooks: initial_sample and transition_sample for one key per particle
    ks = m.random.split(m.random.PRNGKey(2), 16)
    eps = np.stack([orng.normal(k, (d,)) for k in ks.cpu().numpy()]).astype(np.float64)
    x0 = model.initial_sample(0.0, ks)
    np.testing.assert_allclose(x0.cpu().numpy(), opf.init_state(p, eps), rtol=1e-5, atol=1e-5)
    x1 = model.transition_sample(x0, 0.0, 2.25, ks)
    want = opf.transition(p, x0.cpu().numpy().astype(np.float64), eps, 1.5)
    np.testing.assert_allclose(x1.cpu().numpy(), want, rtol=1e-5, atol=1e-5)
    lik = model.likelihood_potential(x1, torch.tensor(y[0], dtype=torch.float32, device="cuda"), 0.0)
    np.testing.assert_allclose(lik.cpu().numpy(), opf.lik_potential(p, x1.cpu().numpy().astype(np.float64), y[0]),
                               rtol=1e-5, atol=1e-4)
    with pytest.raises(NotImplementedError, match="1-D stochastic-volatility"):
        m.ssm.run_particle_filters([model], [y.astype(np.float32)], [key], [n], grad=True, large_n=large_n)


def test_sv_times_through_marginals_and_hooks(m):
    """run_particle_filter_for_marginals passes its t to the kernels; the transition hook with the same
    step scale draws the fused step's particles bit for bit."""
    model = m.ssm.StochasticVolatility(transition_matrix=0.95, transition_covariance=0.09,
                                       likelihood_covariance_diag=0.49)
    y = np.random.default_rng(4).standard_normal((15, 1)).astype(np.float32)
    t = np.cumsum(np.random.default_rng(5).uniform(0.5, 2.5, 15))
    key = m.random.PRNGKey(6)
    a = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), y, t, key, n=256, ess_threshold=0.0,
                                                resampling="systematic")
    b = m.ssm.run_particle_filters([model], [y], [key], [256], ess_threshold=0.0, resampling="systematic",
                                   store=True, ts=[t])[0]
    assert torch.equal(a.value, b.value) and a.log_likelihood.item() == b.log_likelihood.item()
    # never resampling: step 1 propagates the initial particles with the step keys split(split(key)[0], T-1)[0]
    rk = m.random.split(key, 2)[0]
    k1 = m.random.split(m.random.split(rk, 14)[0], 256)
    x1 = model.transition_sample(a.value[0], float(t[0]), float(t[1]), k1)
    assert torch.equal(x1, a.value[1])


def _run_both(m, models, ys, keys, ns, thr, resampling, grad=False):
    kw = dict(ess_threshold=thr, resampling=resampling, store=True, grad=grad)
    big = m.ssm.run_particle_filters(models, ys, keys, ns, large_n=1, **kw)  # every filter chunked
    one = m.ssm.run_particle_filters(models, ys, keys, ns, large_n=1 << 40, **kw)  # a workgroup per filter
    return big, one


@pytest.mark.parametrize("resampling,thr", [("systematic", 0.7), ("multinomial", 0.5), ("systematic", 1.0)])
def test_chunked_path_matches_workgroup_path_sv(m, resampling, thr):
    """Filters on the chunked many-workgroup path draw the same ancestors and particles as on the
    one-workgroup path (same weights, key and quantisation); log-likelihoods agree to FP64 rounding of
    the different summation order. Sizes straddle the 256-particle chunk."""
    from mocat_amd import datasets
    models, ys, _ = datasets.sv_batch_config(4, n_models=7, n_min=64, n_max=3000, t_min=3, t_max=70)
    ns = [1, 255, 256, 257, 700, 2049, 3000]
    keys = [k for k in m.random.split(m.random.PRNGKey(5), 7)]
    big, one = _run_both(m, models, ys, keys, ns, thr, resampling)
    for i in range(7):
        np.testing.assert_array_equal(big[i].ancestors.cpu().numpy(), one[i].ancestors.cpu().numpy())
        np.testing.assert_array_equal(big[i].value.cpu().numpy(), one[i].value.cpu().numpy())
        np.testing.assert_array_equal(big[i].final_value.cpu().numpy(), one[i].final_value.cpu().numpy())
        np.testing.assert_allclose(big[i].ess.cpu().numpy(), one[i].ess.cpu().numpy(), rtol=1e-6)
        a, b = big[i].log_likelihood.item(), one[i].log_likelihood.item()
        assert abs(a - b) <= 1e-10 * (1 + abs(b)), (i, a, b)


def test_chunked_path_lgssm_same_path_as_fp64_oracle(m):
    """A 5000-particle linear-Gaussian filter (20 chunks) on the chunked path vs the FP64 oracle
    replaying its ancestors, and against the one-workgroup path."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(3, 4, 25)
    n = 5000
    yn = y.cpu().numpy()
    key = m.random.PRNGKey(8)
    big, one = _run_both(m, [model], [yn], [key], [n], 0.5, "systematic")
    r = big[0]
    ess = r.ess.cpu().numpy()
    flags = np.zeros(25, bool)
    flags[1:] = ess[:-1] < 0.5 * n
    o = opf.run_pf(_lgssm_np(model), yn, key.cpu().numpy(), n, 0.5, "systematic", ancestors=r.ancestors.cpu().numpy(),
                   resample_flags=flags)
    ll = r.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]), (ll, o["log_likelihood"])
    np.testing.assert_array_equal(r.ancestors.cpu().numpy(), one[0].ancestors.cpu().numpy())
    assert abs(ll - one[0].log_likelihood.item()) <= 1e-10 * abs(ll)


def test_chunked_path_categorical_and_gradient(m):
    """mocat's Gumbel-max resampler and the SV forward-tangent gradient on the chunked path agree with
    the one-workgroup path."""
    from mocat_amd import datasets
    models, ys, _ = datasets.sv_batch_config(9, n_models=3, n_min=64, n_max=600, t_min=5, t_max=30)
    ns = [300, 513, 64]
    keys = [k for k in m.random.split(m.random.PRNGKey(2), 3)]
    big, one = _run_both(m, models, ys, keys, ns, 0.8, "categorical")
    for i in range(3):
        np.testing.assert_array_equal(big[i].ancestors.cpu().numpy(), one[i].ancestors.cpu().numpy())
    big, one = _run_both(m, models, ys, keys, ns, 0.8, "systematic", grad=True)
    for i in range(3):
        np.testing.assert_allclose(big[i].log_likelihood_grad.cpu().numpy(), one[i].log_likelihood_grad.cpu().numpy(),
                                   rtol=1e-9, atol=1e-9)
        T = len(ys[i])
        ess = big[i].ess.cpu().numpy()
        flags = np.zeros(T, bool)
        flags[1:] = ess[:-1] < 0.8 * ns[i]
        og = opf.sv_loglik_grad(models[i].phi, models[i].sigma, models[i].beta, ys[i][:, 0], keys[i].cpu().numpy(),
                                ns[i], big[i].ancestors.cpu().numpy(), flags)
        np.testing.assert_allclose(big[i].log_likelihood_grad.cpu().numpy(), og, rtol=2e-3,
                                   atol=2e-3 * (1 + np.abs(og).max()))


@pytest.mark.parametrize("resampling", ["systematic", "multinomial"])
def test_chunked_path_staged_search_limits(m, resampling):
    """Filters at the edge of the staged two-level ancestor search (2^20 particles: 4096 chunk-end
    sums, the most staged) and just past it (2^20 + 1: the plain search in global memory), batched
    with a small filter in one launch, against the one-workgroup path: same ancestors and particles."""
    from mocat_amd import datasets
    models, ys, _ = datasets.sv_batch_config(6, n_models=3, n_min=64, n_max=3000, t_min=6, t_max=6)
    ns = [1 << 20, (1 << 20) + 1, 3000]
    keys = [k for k in m.random.split(m.random.PRNGKey(9), 3)]
    big, one = _run_both(m, models, ys, keys, ns, 1.0, resampling)
    for i in range(3):
        np.testing.assert_array_equal(big[i].ancestors.cpu().numpy(), one[i].ancestors.cpu().numpy())
        np.testing.assert_array_equal(big[i].final_value.cpu().numpy(), one[i].final_value.cpu().numpy())
        a, b = big[i].log_likelihood.item(), one[i].log_likelihood.item()
        assert abs(a - b) <= 1e-10 * (1 + abs(b)), (i, a, b)


@pytest.mark.parametrize("large_n", [None, 1])
def test_single_observation_and_two_steps(m, large_n):
    """T = 1 (initial weighting only) and T = 2 on both paths against the oracle: log Z_0 =
    lse(lw_0) - log n, and one resampling step."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(2, 2, 2)
    p = _lgssm_np(model)
    yn = y.cpu().numpy()
    for T in (1, 2):
        r = m.ssm.run_particle_filters([model], [yn[:T]], [m.random.PRNGKey(4)], [4096], 1.0, "systematic",
                                       store=True, large_n=large_n)[0]
        o = opf.run_pf(p, yn[:T], orng.PRNGKey(4), 4096, 1.0, "systematic",
                       ancestors=r.ancestors.cpu().numpy(), resample_flags=np.arange(T) > 0)
        assert r.ess.shape == (T,)
        assert abs(r.log_likelihood.item() - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]) + 1e-6
        np.testing.assert_allclose(r.final_value.cpu().numpy(), o["value"][-1], rtol=1e-4, atol=1e-4)


def test_invalid_sizes_raise(m):
    from mocat_amd import datasets
    from mocat_amd._lib import MocatKernelError
    model, y, t = datasets.lgssm_config(0, 4, 5)
    with pytest.raises((MocatKernelError, ValueError)):
        m.ssm.run_particle_filters([model], [y.cpu().numpy()], [m.random.PRNGKey(0)], [0], 1.0, "systematic")


@pytest.mark.parametrize("n,optimal", [(700, False), (5000, False), (1500, True), (6000, True)])
def test_resume_from_initial_sample_repeats_the_run(m, n, optimal):
    """initial_sample: the step keys are split(key, len(y)) directly, so resuming from a run's t = 0
    sample with that run's first split key repeats its steps 1..T-1 bit for bit, on the one-workgroup
    path (n < 4096) and the chunked path; the resumed log-likelihood lacks only the t = 0 term."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(4, 4, 15)
    pf = m.ssm.OptimalNonLinearGaussianParticleFilter() if optimal else m.ssm.BootstrapFilter()
    key = m.random.PRNGKey(9)
    kw = dict(ess_threshold=0.7, resampling="systematic")
    full = m.ssm.run_particle_filter_for_marginals(model, pf, y, t, key, n=n, **kw)
    first = full[0:1]
    res = m.ssm.run_particle_filter_for_marginals(model, pf, y[1:], t[1:], m.random.split(key, 2)[0],
                                                  initial_sample=first, **kw)
    assert res.value.shape == full.value.shape
    np.testing.assert_array_equal(res.value.cpu().numpy(), full.value.cpu().numpy())
    np.testing.assert_array_equal(res.log_weight.cpu().numpy(), full.log_weight.cpu().numpy())
    np.testing.assert_array_equal(res.ancestors.cpu().numpy(), full.ancestors.cpu().numpy())
    np.testing.assert_allclose(res.ess.cpu().numpy(), full.ess.cpu().numpy(), rtol=1e-6)
    lw0 = full.log_weight[0].double().cpu().numpy()
    l0 = opf.lse(lw0) - np.log(n)
    a, b = res.log_likelihood.item(), full.log_likelihood.item() - l0
    assert abs(a - b) <= 1e-9 * (1 + abs(b)), (a, b)
    # and the oracle, replaying the resumed run's ancestors from the same sample, agrees
    if not optimal:
        ess = res.ess.cpu().numpy()
        flags = np.zeros(15, bool)
        flags[1:] = ess[:-1] < 0.7 * n
        o = opf.run_pf(_lgssm_np(model), y.cpu().numpy(), m.random.split(key, 2)[0].cpu().numpy(), n, 0.7,
                       "systematic", ancestors=res.ancestors.cpu().numpy(), resample_flags=flags,
                       initial=(full.value[0].cpu().numpy(), lw0))
        np.testing.assert_allclose(res.value[-1].cpu().numpy(), o["value"][-1], rtol=1e-3, atol=1e-3)
        assert abs(a - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]) + 1e-4


@pytest.mark.parametrize("n,optimal", [(600, False), (5000, False), (800, True), (4500, True)])
def test_one_step_functions_repeat_the_run(m, n, optimal):
    """initiate_particles(split(key)[1]) then propagate_particle_filter with the run's step keys
    split(split(key)[0], T-1)[t-1] reproduce run_particle_filter_for_marginals with mocat's categorical
    resampling bit for bit (resample_full=False keeps the history as the run does); with
    resample_full=True the history is gathered by the same ancestors."""
    from mocat_amd import datasets
    T = 8
    model, y, t = datasets.lgssm_config(6, 4, T)
    pf = m.ssm.OptimalNonLinearGaussianParticleFilter() if optimal else m.ssm.BootstrapFilter()
    key = m.random.PRNGKey(13)
    full = m.ssm.run_particle_filter_for_marginals(model, pf, y, t, key, n=n, ess_threshold=0.6,
                                                   resampling="categorical")
    rkey, sub = m.random.split(key, 2)
    p = m.ssm.initiate_particles(model, pf, n, sub, y[0], float(t[0]))
    np.testing.assert_array_equal(p.value[0].cpu().numpy(), full.value[0].cpu().numpy())
    np.testing.assert_array_equal(p.log_weight[0].cpu().numpy(), full.log_weight[0].cpu().numpy())
    int_keys = m.random.split(rkey, T - 1)
    q = p
    resampled = 0
    for k in range(1, T):
        before = q.value.clone()
        q = m.ssm.propagate_particle_filter(model, pf, q, y[k], float(t[k]), int_keys[k - 1], ess_threshold=0.6,
                                            resample_full=True)
        p = m.ssm.propagate_particle_filter(model, pf, p, y[k], float(t[k]), int_keys[k - 1], ess_threshold=0.6,
                                            resample_full=False)
        np.testing.assert_array_equal(p.value[k].cpu().numpy(), full.value[k].cpu().numpy())
        np.testing.assert_array_equal(p.log_weight[k].cpu().numpy(), full.log_weight[k].cpu().numpy())
        np.testing.assert_array_equal(q.value[k].cpu().numpy(), full.value[k].cpu().numpy())
        anc = full.ancestors[k].long()
        if not torch.equal(anc.cpu(), torch.arange(n)):
            resampled += 1
            np.testing.assert_array_equal(q.value[:k].cpu().numpy(), before[:, anc].cpu().numpy())
    assert p.value.shape == (T, n, 4) and p.ess.shape == (T,) and p.t.shape == (T,)
    np.testing.assert_allclose(p.ess.cpu().numpy(), full.ess.cpu().numpy(), rtol=1e-6)
    assert resampled > 0


def test_initiate_particles_without_observation(m):
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, 3)
    p = m.ssm.initiate_particles(model, m.ssm.BootstrapFilter(), 300, m.random.PRNGKey(2))
    assert p.value.shape == (1, 300, 4) and float(p.log_weight.abs().max()) == 0.0 and float(p.ess[0]) == 300.0
    with pytest.raises(ValueError):
        m.ssm.initiate_particles(model, m.ssm.OptimalNonLinearGaussianParticleFilter(), 300, m.random.PRNGKey(2))


def test_kalman_filter_and_smoother_match_the_oracle(m):
    """mocat's (means, covs) return of the Kalman filter, its log-likelihood attribute, and the RTS
    smoother against oracle.kalman (itself equal to the brute-force joint-Gaussian conditional)."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(2, 4, 12)
    out = m.ssm.run_kalman_filter_for_marginals(model, y, t)
    mus, covs = out
    yn = y.cpu().numpy().astype(np.float64)
    args = (model.initial_mean, model.initial_covariance, model.transition_matrix, model.transition_covariance,
            model.likelihood_matrix, model.likelihood_covariance, yn)
    ll, fm, fc = okal.loglik(*args)
    np.testing.assert_allclose(mus.cpu().numpy(), fm, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(covs.cpu().numpy(), fc, rtol=1e-10, atol=1e-12)
    assert abs(out.log_likelihood.item() - ll) < 1e-9 * abs(ll)
    sm, sc = m.ssm.run_kalman_smoother_for_marginals(model, y, t)
    om, oc = okal.smoother(*args)
    np.testing.assert_allclose(sm.cpu().numpy(), om, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(sc.cpu().numpy(), oc, rtol=1e-9, atol=1e-12)
    sm2, sc2, lag = m.ssm.run_kalman_smoother_for_marginals_and_lag_ones(model, y, t, filter_output=(mus, covs))
    assert torch.equal(sm2, sm) and lag.shape == (11, 4, 4)


def test_chunked_path_wide_ancestor_windows(m):
    """Observation outliers concentrate the weight on a few particles far apart in index, so some
    chunk's systematic targets switch between distant sources: its ancestor window exceeds the LDS
    stage and the search runs over global prefix sums. Ancestors and particles must still equal the
    one-workgroup path's."""
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(5, 4, 10)
    yo = y.cpu().numpy().copy()
    yo[2] += 40.0
    yo[5] -= 60.0
    n = 9000
    big, one = _run_both(m, [model], [yo], [m.random.PRNGKey(21)], [n], 1.0, "systematic")
    anc = big[0].ancestors.cpu().numpy()
    np.testing.assert_array_equal(anc, one[0].ancestors.cpu().numpy())
    np.testing.assert_array_equal(big[0].value.cpu().numpy(), one[0].value.cpu().numpy())
    # an outlier step resamples onto few sources spread over the index range (the -60 outlier may also
    # leave a single source, the fully degenerate case)
    spread = []
    for k in (2, 5):
        src = np.unique(anc[k + 1])
        spread.append((len(src), int(src.max() - src.min())))
    assert any(c < n // 10 and w > 2048 for c, w in spread), spread


def _pf_outputs_equal(a, b):
    for key in ("log_likelihood", "ess", "final_value", "final_log_weight", "value", "log_weight", "ancestors",
                "log_likelihood_grad"):
        if hasattr(a, key):
            assert torch.equal(getattr(a, key), getattr(b, key)), key


@pytest.mark.parametrize("case", ["sv", "sv_grad", "sv3_times", "lgssm_multinomial", "lgssm_categorical",
                                  "optimal"])
def test_persistent_chunked_kernel_repeats_the_stream_launches(m, case):
    """launch "persistent" (one kernel for all steps, grid barriers between the phases) against the
    three launches per step, every filter on the chunked path (large_n = 1), ragged n and T: every
    output bit-identical (log-likelihoods, ESS, stored particles, weights and ancestors, gradients)."""
    from mocat_amd import datasets
    rs = np.random.default_rng(len(case))
    ns = [1, 700, 1024, 1025, 3000, 9000, 20000]
    Ts = [1, 2, 9, 31, 17, 40, 25]
    grad, pf, resampling, ts = False, None, "systematic", None
    if case.startswith("sv"):
        d = 3 if case == "sv3_times" else 1
        model = (m.ssm.StochasticVolatility(3, np.zeros(3), np.eye(3), 0.9 * np.eye(3), 0.05 * np.eye(3), np.ones(3))
                 if d == 3 else m.ssm.StochasticVolatility(1, 0.0, 1.0, 0.95, 0.04, 0.6))
        p = (opf.svd_params(np.zeros(3), np.eye(3), 0.9 * np.eye(3), 0.05 * np.eye(3), np.ones(3)) if d == 3
             else opf.sv_params(0.95, 0.2, np.sqrt(0.6)))
        ys = [opf.simulate(p, T, orng.PRNGKey(T))[1].astype(np.float32) for T in Ts]
        grad = case == "sv_grad"
        if d == 3:
            ts = [np.cumsum(rs.uniform(0.2, 2.0, T)) for T in Ts]
    else:
        model, y, _ = datasets.lgssm_config(0, 4, 40)
        ys = [y.cpu().numpy()[:T] for T in Ts]
        resampling = {"lgssm_multinomial": "multinomial", "lgssm_categorical": "categorical"}.get(case, "systematic")
        if case == "optimal":
            pf = m.ssm.OptimalNonLinearGaussianParticleFilter
    if resampling == "categorical":
        ns = [v for v in ns if v * v < 0xFFFFFFFF]
        ys = ys[:len(ns)]
    keys = [k for k in m.random.split(m.random.PRNGKey(4), len(ns))]
    models = [model] * len(ns)
    thr = 0.7
    run = lambda launch: m.ssm.run_particle_filters(models, ys, keys, ns, thr, resampling, store=True, grad=grad,
                                                    particle_filter=pf, ts=ts,
                                                    schedule={"large_n": 1, "launch": launch})
    base, pers = run("stream"), run("persistent")
    for a, b in zip(base, pers):
        _pf_outputs_equal(a, b)
        assert np.isfinite(b.log_likelihood.item())
    # and the stream launches repeat the one-workgroup path's filters to FP64 summation order
    small = m.ssm.run_particle_filters(models, ys, keys, ns, thr, resampling, store=True, grad=grad,
                                       particle_filter=pf, ts=ts, large_n=1 << 40)
    for a, b in zip(small, pers):
        x, y = a.log_likelihood.item(), b.log_likelihood.item()
        assert abs(x - y) <= 1e-9 * (1 + abs(y)), (x, y)


def test_persistent_kernel_barrier_gives_up_cleanly(m, monkeypatch):
    """The persistent kernel's grid barriers wait a bounded number of polls. With the limit forced to
    zero from the host every workgroup gives up at the first barrier and leaves the kernel, and the call
    returns NaN log-likelihoods (and gradients); the next call with the default limit is unaffected."""
    from mocat_amd import datasets
    models, ys, ns = datasets.sv_batch_config(3, n_models=4, n_min=3000, n_max=9000, t_min=10, t_max=20)
    keys = [k for k in m.random.split(m.random.PRNGKey(6), len(models))]
    run = lambda launch: m.ssm.run_particle_filters(models, ys, keys, ns, 1.0, "systematic", grad=True,
                                                    schedule={"large_n": 1, "launch": launch})
    base = run("stream")
    monkeypatch.setenv("MOCAT_PF_BARRIER_POLLS", "0")
    bad = run("persistent")
    assert all(np.isnan(r.log_likelihood.item()) for r in bad)
    assert all(torch.isnan(r.log_likelihood_grad).all() for r in bad)
    monkeypatch.delenv("MOCAT_PF_BARRIER_POLLS")
    for a, b in zip(base, run("persistent")):
        _pf_outputs_equal(a, b)


@pytest.mark.parametrize("grad", [False, True])
def test_graph_launch_and_tuned_schedule_give_the_same_results(m, grad, tmp_path, monkeypatch):
    # the chunked path's per-step launches captured into one HIP graph, and the autotuner's choice of
    # threshold and launch mode: every schedule returns the same log-likelihoods, ESS and gradients
    from mocat_amd import datasets
    from mocat_amd.ssm import schedule as sched
    models, ys, ns = datasets.sv_batch_config(3, n_models=12, n_min=300, n_max=20000, t_min=20, t_max=120)
    keys = [k for k in m.random.split(m.random.PRNGKey(5), len(models))]
    run = lambda **kw: m.ssm.run_particle_filters(models, ys, keys, ns, 1.0, "systematic", grad=grad, **kw)
    base = run(large_n=4096)
    assert any(n >= 4096 for n in ns) and any(n < 4096 for n in ns)
    monkeypatch.setenv("MOCAT_AMD_SCHEDULE_CACHE", str(tmp_path / "sched.json"))
    sched._MEMO.clear()
    outs = [run(schedule={"large_n": 4096, "graph": True}), run(schedule={"large_n": 4096, "launch": "persistent"}),
            run(schedule={"large_n": 1024, "graph": True}), run(schedule="auto")]
    rec = sched.lookup(sched.signature(models, ys, ns, "systematic", grad, False, None))
    assert rec is not None and rec["large_n"] in sched.CANDIDATE_LARGE_N and rec["launch"] in sched.CANDIDATE_LAUNCH
    assert len(rec["timings_ms"]) == len(sched.CANDIDATE_LARGE_N) * len(sched.CANDIDATE_LAUNCH)
    assert (tmp_path / "sched.json").exists()
    # the same threshold with the graph or the persistent kernel: bit-identical; another threshold moves
    # filters between the two paths, whose FP64 sums differ only in their order (as
    # test_chunked_path_matches_workgroup_path_sv)
    for k, out in enumerate(outs):
        for a, b in zip(base, out):
            if k <= 1:
                assert a.log_likelihood.item() == b.log_likelihood.item()
                assert torch.equal(a.ess, b.ess)
                if grad:
                    assert torch.equal(a.log_likelihood_grad, b.log_likelihood_grad)
            else:
                x, y = a.log_likelihood.item(), b.log_likelihood.item()
                assert abs(x - y) <= 1e-10 * (1 + abs(y))
                np.testing.assert_allclose(a.ess.cpu().numpy(), b.ess.cpu().numpy(), rtol=1e-6)
                if grad:
                    np.testing.assert_allclose(a.log_likelihood_grad.cpu().numpy(),
                                               b.log_likelihood_grad.cpu().numpy(), rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("n,T", [(4096, 600), (5003, 520)])
def test_chunked_sv_long_run_same_path_as_fp64_oracle(m, n, T):
    # the chunked many-workgroup path (filters from 4096 particles) at config-5 lengths: the FP64 oracle
    # replaying the GPU's ancestors gives the same log-likelihood to 1e-4 relative after T steps of FP64
    # accumulation, and the same log-likelihood gradient (fixed-resampling-index estimator)
    from mocat_amd import datasets
    models, ys, _ = datasets.sv_batch_config(8, n_models=1, t_min=T, t_max=T)
    key = m.random.PRNGKey(21)
    r = m.ssm.run_particle_filters(models, ys, [key], [n], ess_threshold=1.0, resampling="systematic", store=True,
                                   grad=True, large_n=4096)[0]
    md = models[0]
    p = opf.sv_params(md.phi, md.sigma, md.beta)
    flags = np.ones(T, bool)
    flags[0] = False
    anc = r.ancestors.cpu().numpy()
    o = opf.run_pf(p, ys[0], key.cpu().numpy(), n, 1.0, "systematic", ancestors=anc, resample_flags=flags)
    ll = r.log_likelihood.item()
    assert abs(ll - o["log_likelihood"]) <= 1e-4 * abs(o["log_likelihood"]), (ll, o["log_likelihood"])
    g = opf.sv_loglik_grad(md.phi, md.sigma, md.beta, ys[0], key.cpu().numpy(), n, anc, flags)
    np.testing.assert_allclose(r.log_likelihood_grad.cpu().numpy(), g, rtol=2e-3, atol=2e-3 * np.abs(g).max())
