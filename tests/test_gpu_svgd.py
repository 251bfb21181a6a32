"""GPU SVGD pieces vs the FP64 oracle: GMM potential/adjoint, pairwise distances, exact median,
one Stein step with adagrad, and an end-to-end run on a Gaussian target (mocat's SVGD test pattern)."""
import math

import numpy as np
import pytest
import torch

from oracle import svgd as osv

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def _gmm_case(seed, K=4, d=100):
    rs = np.random.default_rng(seed)
    return rs.normal(0, 3, (K, d)), rs.uniform(0.5, 2, (K, d)), rs.dirichlet(np.ones(K))


@pytest.mark.parametrize("conv", ["textbook", "mocat"])
def test_gmm_potential_grad(m, conv):
    mu, var, w = _gmm_case(0)
    sc = m.GaussianMixture(mu, var, w, det_convention=conv)
    x = np.random.default_rng(1).normal(0, 3, (300, 100)).astype(np.float32)
    _, pot, _, grad = sc.evaluate(torch.tensor(x, device="cuda"))
    opot, ograd = osv.gmm_potential_grad(x, mu, var, w, conv)
    np.testing.assert_allclose(pot.cpu().numpy(), opot, rtol=2e-5, atol=1e-3)
    np.testing.assert_allclose(grad.cpu().numpy(), ograd, rtol=1e-4, atol=1e-3)


def _spd(rs, K, d):
    A = rs.normal(0, 1, (K, d, d))
    return np.einsum("kij,klj->kil", A, A) / d + 0.5 * np.eye(d)[None]


@pytest.mark.parametrize("conv", ["textbook", "mocat"])
@pytest.mark.parametrize("n,d,K", [(1000, 7, 3), (513, 100, 4), (65, 128, 8), (3, 1, 2)])
def test_gmm_full_covariance_vs_fp64(m, conv, n, d, K):
    # mocat's GaussianMixture with full covariances (toy_examples.py:73-103), both det/quadratic conventions
    rs = np.random.default_rng(d + K)
    mu = rs.normal(0, 1.5, (K, d))
    covs = _spd(rs, K, d)
    w = rs.dirichlet(np.ones(K))
    sc = m.GaussianMixture(mu, covs, w, det_convention=conv)
    assert sc.full == (d > 1)  # 1 x 1 covariances are diagonal: the diagonal kernel
    x = rs.normal(0, 1.5, (n, d)).astype(np.float32)
    _, pot, _, grad = sc.evaluate(torch.tensor(x, device="cuda"))
    opot, ograd = osv.gmm_full_potential_grad(x, mu, covs, w, conv)
    scale = np.abs(opot) + d
    assert np.max(np.abs(pot.cpu().numpy() - opot) / scale) < 1e-5
    gscale = np.abs(ograd).max(1, keepdims=True) + 1e-3
    assert np.max(np.abs(grad.cpu().numpy() - ograd) / gscale) < 5e-5


def test_gmm_full_covariance_samplers_follow_user_scenario(m):
    # the full-covariance mixture through HMC (the iteration in pieces) and SVGD (stepwise path) equals
    # the same potential written as a user scenario in torch, on the same keys
    rs = np.random.default_rng(5)
    K, d = 3, 6
    mu = rs.normal(0, 2, (K, d))
    covs = _spd(rs, K, d)
    w = rs.dirichlet(np.ones(K))
    fused = m.GaussianMixture(mu, covs, w)
    S, logc, mut = fused.S, fused.logc, fused.means

    class UserMixture(m.Scenario):
        name = "user mixture"
        dim = d

        def prior_potential(self, x, random_key=None):
            return torch.sum(0.0 * x)

        def likelihood_potential(self, x, random_key=None):
            r = x - mut
            return -torch.logsumexp(logc - 0.5 * torch.einsum("ki,kij,kj->k", r, S, r), 0)

    user = UserMixture()
    hmc = [m.run(sc, m.Underdamped(stepsize=0.2, leapfrog_steps=3), 20, m.random.PRNGKey(2), n_chains=64,
                 correction=m.Metropolis()) for sc in (fused, user)]
    same = (hmc[0].value - hmc[1].value).abs().amax(dim=(0, 2)) < 1e-4
    assert same.float().mean().item() >= 0.9
    x0 = torch.tensor(rs.normal(0, 1, (256, d)), dtype=torch.float32, device="cuda")
    sv = [m.run(sc, m.SVGD(stepsize=0.1, max_iter=4, bandwidth_update="median"), 256, m.random.PRNGKey(1),
                initial_state=m.cdict(value=x0.clone()), stack=("value",)).value for sc in (fused, user)]
    torch.testing.assert_close(sv[0], sv[1], rtol=1e-4, atol=1e-4)


def test_gmm_reference_kats_on_gpu(m):
    sc = m.GaussianMixture(np.arange(15.0).reshape(5, 3))
    x = torch.tensor([[0.0, 0, 0], [2.0, 2, 2]], device="cuda")
    pot, grad = sc.potential_and_grad(x)
    np.testing.assert_allclose(pot.cpu().numpy(), [6.8662534, 6.8552055], rtol=1e-6)
    np.testing.assert_allclose(grad.cpu().numpy(), [[0.0, -1.0, -2.0], [1.967039, 0.967039, -0.032961]], atol=2e-6)


@pytest.mark.parametrize("n,d", [(4096, 100), (300, 7), (33, 64), (1, 5)])
def test_pairwise_sqdist_vs_fp64(m, n, d):
    x = np.random.default_rng(n).normal(0, 3, (n, d)).astype(np.float32) + 5.0
    d2 = m.kernels.pairwise_sqdist(torch.tensor(x, device="cuda")).cpu().numpy().astype(np.float64)
    if n <= 300:
        ref = osv.distance_matrix(x) ** 2
    else:
        xs = x.astype(np.float64)
        sq = (xs * xs).sum(1)
        ref = np.maximum(sq[:, None] + sq[None, :] - 2 * xs @ xs.T, 0)
        np.fill_diagonal(ref, 0)
    assert np.all(np.diag(d2) == 0)
    scale = ref.max() + 1e-30
    assert np.max(np.abs(d2 - ref)) / scale < 2e-5


@pytest.mark.parametrize("m_", [1, 2, 1000, 1 << 20])
def test_select_exact(m, m_):
    v = np.random.default_rng(m_).exponential(size=m_).astype(np.float32)
    v[: m_ // 10] = 0.0
    r0, r1 = (m_ - 1) // 2, m_ // 2
    got = m.kernels.select_pair(torch.tensor(v, device="cuda"), r0, r1).cpu().numpy()
    s = np.sort(v)
    assert got[0] == s[r0] and got[1] == s[r1]


@pytest.mark.parametrize("case", ["extremes", "thirds", "constant", "two_values", "narrow"])
def test_select_exact_rank_pairs(m, case):
    """Ranks whose prefixes part at the first byte, and inputs whose values share most bytes (every
    lane of a wave on one digit), against a full sort."""
    rs = np.random.default_rng(5)
    m_ = 300_001
    if case == "constant":
        v = np.full(m_, 3.5, np.float32)
    elif case == "two_values":
        v = np.where(rs.random(m_) < 0.5, 1.0, 1e6).astype(np.float32)
    elif case == "narrow":
        v = (100.0 + rs.random(m_) * 1e-3).astype(np.float32)
    else:
        v = rs.lognormal(0, 4, m_).astype(np.float32)
    r0, r1 = {"extremes": (0, m_ - 1), "thirds": (m_ // 3, 2 * m_ // 3)}.get(case, (m_ // 2 - 1, m_ // 2 + 7))
    got = m.kernels.select_pair(torch.tensor(v, device="cuda"), r0, r1).cpu().numpy()
    s = np.sort(v)
    assert got[0] == s[r0] and got[1] == s[r1]


def test_median_and_mean_bandwidth(m):
    x = np.random.default_rng(3).normal(0, 2, (700, 30)).astype(np.float32)
    xt = torch.tensor(x, device="cuda")
    np.testing.assert_allclose(m.kernels.median_bandwidth_update(xt).item(), osv.median_bandwidth(x), rtol=2e-5)
    np.testing.assert_allclose(m.kernels.mean_bandwidth_update(xt).item(), osv.mean_bandwidth(x), rtol=2e-5)


@pytest.mark.parametrize("n,d,js", [(4096, 100, None), (500, 10, 4), (130, 65, None), (33, 3, 1), (300, 128, 64)])
def test_svgd_step_vs_fp64(m, n, d, js):
    """Ragged particle tiles (n not a multiple of 32 or 128), both packed widths, and j splits from
    one to more than there are j blocks (empty ranges contribute zero partial sums)."""
    rs = np.random.default_rng(d)
    x = rs.normal(0, 2, (n, d)).astype(np.float32)
    g = rs.normal(0, 1, (n, d)).astype(np.float32)
    h = float(osv.median_bandwidth(x)) if n <= 500 else 3.0
    smp = m.SVGD(stepsize=0.01, bandwidth_update=None, j_splits=js)
    st = m.cdict(value=torch.tensor(x, device="cuda"), grad_potential=torch.tensor(g, device="cuda"))
    ex = m.cdict(random_key=m.random.PRNGKey(0), iter=0, bandwidth=h, g_sq=torch.zeros((n, d), device="cuda"),
                 m=torch.zeros((n, d), device="cuda"))
    sc = m.DiagonalGaussian(np.zeros(d), np.ones(d), prior_variance=math.inf)
    phi = smp.step(st, ex, sc).cpu().numpy().astype(np.float64)
    ophi = osv.phi_gemm(x, g, h)
    err = np.max(np.abs(phi - ophi)) / np.max(np.abs(ophi))
    assert err < 1e-4, err
    ox, gsq, mm = osv.adagrad(x.astype(np.float64), np.zeros((n, d)), np.zeros((n, d)), -ophi, 0.01)
    np.testing.assert_allclose(st.value.cpu().numpy(), ox, atol=1e-4)


def test_svgd_run_gaussian_moments(m):
    # mocat/src/tests/test_transport.py:65-89 pattern: moments of the final ensemble
    sc = m.DiagonalGaussian(np.zeros(2), np.array([1.0, 2.0]), prior_variance=math.inf)
    for upd in ("median", "mean"):
        chain = m.run(sc, m.SVGD(stepsize=1.0, max_iter=500, bandwidth_update=upd), 256, m.random.PRNGKey(0))
        x = chain.final_state.value.cpu().numpy()
        np.testing.assert_allclose(x.mean(0), [0, 0], atol=0.1)
        np.testing.assert_allclose(np.var(x, 0), [1.0, 2.0], rtol=0.2)


@pytest.mark.parametrize("n", [3, 4, 5, 33, 130, 1000])
def test_median_over_pairs_exact(m, n):
    """The pair-only selection returns exactly the median order statistics of the full n x n multiset
    (n diagonal zeros and each pair twice) built from the same GPU distances."""
    rs = np.random.default_rng(n)
    x = rs.normal(0, 1.5, (n, 7)).astype(np.float32)
    x[n // 2] = x[0]  # one zero-distance pair
    d2 = m.kernels._sqdist_upper(torch.tensor(x, device="cuda")).cpu().numpy()
    iu = np.triu_indices(n, 1)
    full = np.sort(np.concatenate([np.zeros(n, np.float32), d2[iu], d2[iu]]))
    mm = n * n
    want = 0.5 * (np.sqrt(np.float64(full[(mm - 1) // 2])) + np.sqrt(np.float64(full[mm // 2])))
    want = np.float32(want / np.sqrt(2 * np.log(n)))
    got = m.kernels.median_bandwidth_update(torch.tensor(x, device="cuda")).item()
    assert got == want
    mean = np.float32(2 * np.sqrt(d2[iu].astype(np.float64)).sum() / (n * n) / np.sqrt(2 * np.log(n)))
    assert m.kernels.mean_bandwidth_update(torch.tensor(x, device="cuda")).item() == pytest.approx(float(mean), rel=1e-6)


def test_select_pairs_constant_and_ties(m):
    """Every pair at one distance, and two clusters (ties across the whole top byte)."""
    n = 300
    x = np.zeros((n, 3), np.float32)
    x[:, 0] = np.arange(n) % 2  # pair distances are 0 or 1
    got = m.kernels.median_bandwidth_update(torch.tensor(x, device="cuda")).item()
    full = np.sort(np.concatenate([np.zeros(n), (np.subtract.outer(x[:, 0], x[:, 0]) ** 2)[np.triu_indices(n, 1)].repeat(2)]))
    mm = n * n
    want = np.float32(0.5 * (np.sqrt(full[(mm - 1) // 2]) + np.sqrt(full[mm // 2])) / np.sqrt(2 * np.log(n)))
    assert got == want


def test_svgd_step_device_bandwidth_matches_host(m):
    """The kernel reads the bandwidth from device memory with the same arithmetic as from the host."""
    rs = np.random.default_rng(1)
    n, d = 700, 30
    x = rs.normal(0, 2, (n, d)).astype(np.float32)
    g = rs.normal(0, 1, (n, d)).astype(np.float32)
    outs = []
    for bw in (2.5, torch.tensor(2.5, device="cuda")):
        smp = m.SVGD(stepsize=0.01, bandwidth_update=None)
        st = m.cdict(value=torch.tensor(x, device="cuda"), grad_potential=torch.tensor(g, device="cuda"))
        ex = m.cdict(random_key=m.random.PRNGKey(0), iter=0, bandwidth=bw, g_sq=torch.zeros((n, d), device="cuda"),
                     m=torch.zeros((n, d), device="cuda"))
        sc = m.DiagonalGaussian(np.zeros(d), np.ones(d), prior_variance=math.inf)
        outs.append((smp.step(st, ex, sc).cpu().numpy(), st.value.cpu().numpy()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("upd", ["median", "mean", None])
def test_fused_gmm_run_matches_stepwise(m, upd):
    """The single-call SVGD run on a Gaussian mixture reproduces repeated step() calls: particles,
    adagrad state, bandwidth trace and key (median and fixed bandwidths bit for bit; the mean's FP64
    sum order differs)."""
    rs = np.random.default_rng(2)
    d, n, iters = 6, 300, 12
    sc = m.GaussianMixture(rs.normal(0, 2, (3, d)), rs.uniform(0.5, 2, (3, d)), rs.dirichlet(np.ones(3)))
    kw = dict(stepsize=0.05, max_iter=iters, bandwidth_update=upd)
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(4))
    smp = m.SVGD(**kw)
    ref = m.run(sc, smp, n, m.random.PRNGKey(4), stack=("value",))  # stack forces the stepwise loop
    fx, rx = fused.final_state.value.cpu().numpy(), ref.final_state.value.cpu().numpy()
    if upd == "mean":
        np.testing.assert_allclose(fx, rx, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy(), rtol=1e-6)
    else:
        np.testing.assert_array_equal(fx, rx)
        np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())
        np.testing.assert_array_equal(fused.final_extra.g_sq.cpu().numpy(), ref.final_extra.g_sq.cpu().numpy())
    np.testing.assert_array_equal(fused.final_extra.random_key.cpu().numpy(), ref.final_extra.random_key.cpu().numpy())


def test_fused_median_all_selection_routes(m):
    """n = 3000: the first selection runs over all pairs, the second on a wide bracket (many
    candidates, multi-workgroup passes), later ones on few candidates (one workgroup in LDS); every
    bandwidth equals the stepwise full selection bit for bit."""
    rs = np.random.default_rng(11)
    d, n, iters = 20, 3000, 8
    sc = m.GaussianMixture(rs.normal(0, 1, (3, d)), rs.uniform(0.5, 2, (3, d)), rs.dirichlet(np.ones(3)))
    kw = dict(stepsize=0.01, max_iter=iters, bandwidth_update="median")
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(3))
    ref = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(3), stack=("value",))
    np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())
    np.testing.assert_array_equal(fused.final_state.value.cpu().numpy(), ref.final_state.value.cpu().numpy())


def test_fused_median_bracket_exact_over_many_iterations(m):
    """Small steps keep the median inside the bracket of the previous one, so most selections run on
    the candidate set: the bandwidth trace still equals the stepwise full selection bit for bit."""
    rs = np.random.default_rng(7)
    d, n, iters = 10, 1000, 40
    sc = m.GaussianMixture(rs.normal(0, 1, (2, d)), rs.uniform(0.5, 2, (2, d)), np.array([0.3, 0.7]))
    kw = dict(stepsize=0.005, max_iter=iters, bandwidth_update="median")
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(1))
    ref = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(1), stack=("value",))
    np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())
    np.testing.assert_array_equal(fused.final_state.value.cpu().numpy(), ref.final_state.value.cpu().numpy())


def test_fused_median_barrier_time_out_gives_nan_bandwidth(m, monkeypatch):
    """The digit passes' grid barrier with a zero wait limit, forced from the host: the first selection
    runs over all pairs, every barrier times out and the bandwidth is NaN. Without the limit the same
    run gives the stepwise bandwidths again (the barrier state is reset by the finishing kernel)."""
    rs = np.random.default_rng(11)
    d, n, iters = 20, 3000, 3
    sc = m.GaussianMixture(rs.normal(0, 1, (3, d)), rs.uniform(0.5, 2, (3, d)), rs.dirichlet(np.ones(3)))
    kw = dict(stepsize=0.01, max_iter=iters, bandwidth_update="median")
    monkeypatch.setenv("MOCAT_SVGD_BARRIER_POLLS", "0")
    bad = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(3))
    bws = bad.bandwidth.cpu().numpy()  # the starting bandwidth, then one per iteration
    assert np.isfinite(bws[0]) and np.isnan(bws[1])
    monkeypatch.delenv("MOCAT_SVGD_BARRIER_POLLS")
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(3))
    ref = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(3), stack=("value",))
    np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())


@pytest.mark.parametrize("n", [3, 37, 130])
def test_fused_median_small_ensembles(m, n):
    """Few particles: the selection's grid is n - 1 workgroups (2 at n = 3), its first selection runs
    over all pairs behind grid barriers and later ones on the bracket candidates; every bandwidth and
    the final particles equal the stepwise full selection bit for bit."""
    rs = np.random.default_rng(n)
    d, iters = 7, 12
    sc = m.GaussianMixture(rs.normal(0, 1, (2, d)), rs.uniform(0.5, 2, (2, d)), np.array([0.4, 0.6]))
    kw = dict(stepsize=0.01, max_iter=iters, bandwidth_update="median")
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(2))
    ref = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(2), stack=("value",))
    assert np.isfinite(fused.bandwidth.cpu().numpy()).all()
    np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())
    np.testing.assert_array_equal(fused.final_state.value.cpu().numpy(), ref.final_state.value.cpu().numpy())


def test_fused_median_large_ensemble(m):
    """16384 particles: 8256 distance workgroups add to the striped count-below counters and append
    candidates; the bandwidths and final particles equal the stepwise full selection bit for bit."""
    rs = np.random.default_rng(5)
    d, n, iters = 20, 16384, 4
    sc = m.GaussianMixture(rs.normal(0, 1, (3, d)), rs.uniform(0.5, 2, (3, d)), rs.dirichlet(np.ones(3)))
    kw = dict(stepsize=0.01, max_iter=iters, bandwidth_update="median")
    fused = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(6))
    ref = m.run(sc, m.SVGD(**kw), n, m.random.PRNGKey(6), stack=("value",))
    np.testing.assert_array_equal(fused.bandwidth.cpu().numpy(), ref.bandwidth.cpu().numpy())
    np.testing.assert_array_equal(fused.final_state.value.cpu().numpy(), ref.final_state.value.cpu().numpy())
