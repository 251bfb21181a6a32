"""Pins the SVGD/GMM oracle with mocat's own known answers and a brute-force phi."""
import numpy as np

from oracle import svgd


def test_gaussian_kernel_reference_kat():
    # mocat/src/tests/test_kernels.py:19-21 (bandwidth 1)
    np.testing.assert_allclose(svgd.gaussian_kernel(np.zeros(5), np.zeros(5), 1.0), 1.0)
    np.testing.assert_allclose(svgd.gaussian_kernel(np.zeros(5), np.ones(5), 1.0), 0.082085006, rtol=1e-7)


def test_gmm_reference_kats():
    # 1-D: mocat/src/tests/test_toy_scenarios.py:59-75 (means [0, 1], identity covariances)
    pot, grad = svgd.gmm_potential_grad(np.array([[5.0]]), [[0.0], [1.0]], np.ones((2, 1)), [0.5, 0.5], "mocat")
    np.testing.assert_allclose(pot, [9.601038], rtol=1e-6)
    np.testing.assert_allclose(grad[0], [4.010987], rtol=1e-6)
    # N-D: :86-106 (means arange(15).reshape(5, 3))
    mu = np.arange(15.0).reshape(5, 3)
    for conv in ("mocat", "textbook"):  # identity covariances: both conventions coincide
        pot, grad = svgd.gmm_potential_grad(np.array([[0.0, 0, 0], [2.0, 2, 2]]), mu, np.ones((5, 3)),
                                            np.ones(5) / 5, conv)
        np.testing.assert_allclose(pot, [6.8662534, 6.8552055], rtol=1e-6)
        np.testing.assert_allclose(grad[0], [0.0, -1.0, -2.0], atol=1e-6)
        np.testing.assert_allclose(grad[1], [1.967039, 0.967039, -0.032961], atol=2e-6)


def test_gmm_textbook_is_normalised_mixture_logpdf():
    from scipy.stats import multivariate_normal
    rs = np.random.default_rng(0)
    mu = rs.normal(0, 3, (3, 4))
    var = rs.uniform(0.5, 2, (3, 4))
    w = rs.dirichlet(np.ones(3))
    x = rs.normal(0, 3, (6, 4))
    dens = sum(w[k] * multivariate_normal(mu[k], np.diag(var[k])).pdf(x) for k in range(3))
    pot, _ = svgd.gmm_potential_grad(x, mu, var, w, "textbook")
    np.testing.assert_allclose(pot, -np.log(dens), rtol=1e-10)


def test_phi_gemm_form_equals_pairwise_sum():
    rs = np.random.default_rng(1)
    x = rs.normal(size=(40, 5))
    g = rs.normal(size=(40, 5))
    np.testing.assert_allclose(svgd.phi_gemm(x, g, 0.7), svgd.phi_pairwise(x, g, 0.7), rtol=1e-10, atol=1e-12)


def test_gmm_grad_central_differences():
    rs = np.random.default_rng(2)
    mu, var, w = rs.normal(0, 2, (4, 3)), rs.uniform(0.5, 2, (4, 3)), rs.dirichlet(np.ones(4))
    x = rs.normal(size=(1, 3))
    _, g = svgd.gmm_potential_grad(x, mu, var, w)
    eps = 1e-6
    fd = [(svgd.gmm_potential_grad(x + eps * e, mu, var, w)[0] - svgd.gmm_potential_grad(x - eps * e, mu, var, w)[0])
          / (2 * eps) for e in np.eye(3)]
    np.testing.assert_allclose(g[0], np.array(fd)[:, 0], rtol=1e-6, atol=1e-8)


def test_adagrad_first_step_scale_invariance_and_constant_gradient():
    """jax.example_libraries.optimizers.adagrad (used by mocat/src/transport/svgd.py:11,56): the first
    step moves by step (1 - beta) sign(g) (zero for g = 0); scaling every gradient by c > 0 leaves the
    path unchanged; for a constant gradient m_k = (1 - beta) sign(g) sum_i beta^(k-i) / sqrt(i)."""
    from oracle import svgd as osvgd
    beta, step = 0.9, 0.05
    g = np.array([2.0, -0.5, 0.0])
    x0 = np.array([1.0, 1.0, 1.0])
    x1, gsq, m = osvgd.adagrad(x0, np.zeros(3), np.zeros(3), g, step, beta)
    np.testing.assert_allclose(x1, x0 - step * (1 - beta) * np.sign(g), rtol=0, atol=1e-15)
    rs = np.random.default_rng(0)
    gs = rs.normal(0, 1, (6, 3))
    xa = xb = x0
    sa = sb = (np.zeros(3), np.zeros(3))
    for k in range(6):
        xa, *sa = osvgd.adagrad(xa, sa[0], sa[1], gs[k], step, beta)
        xb, *sb = osvgd.adagrad(xb, sb[0], sb[1], 37.0 * gs[k], step, beta)
    np.testing.assert_allclose(xa, xb, rtol=1e-12)
    x, s = x0, (np.zeros(3), np.zeros(3))
    gc = np.array([3.0, -1.0, 0.25])
    for k in range(1, 8):
        x, *s = osvgd.adagrad(x, s[0], s[1], gc, step, beta)
        want = (1 - beta) * np.sign(gc) * sum(beta ** (k - i) / np.sqrt(i) for i in range(1, k + 1))
        np.testing.assert_allclose(s[1], want, rtol=1e-12)


def _spd(rs, K, d):
    A = rs.normal(0, 1, (K, d, d))
    return np.einsum("kij,klj->kil", A, A) / d + 0.5 * np.eye(d)[None]


def test_gmm_full_reference_kats_and_diagonal_case():
    # mocat's known answers (identity covariances) through the full-covariance form
    pot, grad = svgd.gmm_full_potential_grad(np.array([[5.0]]), [[0.0], [1.0]], np.ones((2, 1, 1)), [0.5, 0.5],
                                             "mocat")
    np.testing.assert_allclose(pot, [9.601038], rtol=1e-6)
    np.testing.assert_allclose(grad[0], [4.010987], rtol=1e-6)
    mu = np.arange(15.0).reshape(5, 3)
    pot, grad = svgd.gmm_full_potential_grad(np.array([[0.0, 0, 0], [2.0, 2, 2]]), mu,
                                             np.repeat(np.eye(3)[None], 5, 0), np.ones(5) / 5, "mocat")
    np.testing.assert_allclose(pot, [6.8662534, 6.8552055], rtol=1e-6)
    np.testing.assert_allclose(grad[1], [1.967039, 0.967039, -0.032961], atol=2e-6)
    # diagonal covariances: the full form reduces to the diagonal one under both conventions
    rs = np.random.default_rng(1)
    mu = rs.normal(0, 2, (3, 5))
    var = rs.uniform(0.3, 3.0, (3, 5))
    w = rs.dirichlet(np.ones(3))
    x = rs.normal(0, 2, (7, 5))
    covs = np.stack([np.diag(v) for v in var])
    for conv in ("mocat", "textbook"):
        p1, g1 = svgd.gmm_full_potential_grad(x, mu, covs, w, conv)
        p0, g0 = svgd.gmm_potential_grad(x, mu, var, w, conv)
        np.testing.assert_allclose(p1, p0, rtol=1e-12)
        np.testing.assert_allclose(g1, g0, rtol=1e-10, atol=1e-12)


def test_gmm_full_textbook_is_normalised_mixture_logpdf():
    from scipy.stats import multivariate_normal
    rs = np.random.default_rng(2)
    K, d = 3, 4
    mu = rs.normal(0, 2, (K, d))
    covs = _spd(rs, K, d)
    w = rs.dirichlet(np.ones(K))
    x = rs.normal(0, 2, (6, d))
    dens = sum(w[k] * multivariate_normal(mu[k], covs[k]).pdf(x) for k in range(K))
    pot, _ = svgd.gmm_full_potential_grad(x, mu, covs, w, "textbook")
    np.testing.assert_allclose(pot, -np.log(dens), rtol=1e-10)


def test_gmm_full_mocat_form_is_mixture_with_precision_LtL():
    # mocat's term w det(P) (2 pi)^(-d/2) exp(-0.5 |L r|^2) is w sqrt(det P) N(x; mu, (L^T L)^-1) because
    # det(L^T L) = det(P): a transposed factor (L L^T = P) fails this for non-diagonal covariances
    from scipy.stats import multivariate_normal
    rs = np.random.default_rng(3)
    K, d = 2, 3
    mu = rs.normal(0, 1, (K, d))
    covs = _spd(rs, K, d)
    w = rs.dirichlet(np.ones(K))
    x = rs.normal(0, 1, (5, d))
    dens = 0.0
    for k in range(K):
        P = np.linalg.inv(covs[k])
        L = np.linalg.cholesky(P)
        dens = dens + w[k] * np.sqrt(np.linalg.det(P)) * multivariate_normal(mu[k], np.linalg.inv(L.T @ L)).pdf(x)
    pot, _ = svgd.gmm_full_potential_grad(x, mu, covs, w, "mocat")
    np.testing.assert_allclose(pot, -np.log(dens), rtol=1e-10)


def test_gmm_full_grad_central_differences():
    rs = np.random.default_rng(4)
    K, d = 3, 4
    mu = rs.normal(0, 1, (K, d))
    covs = _spd(rs, K, d)
    w = rs.dirichlet(np.ones(K))
    x = rs.normal(0, 1, (3, d))
    for conv in ("mocat", "textbook"):
        _, g = svgd.gmm_full_potential_grad(x, mu, covs, w, conv)
        eps = 1e-6
        for j in range(d):
            e = np.zeros(d)
            e[j] = eps
            fd = (svgd.gmm_full_potential_grad(x + e, mu, covs, w, conv)[0]
                  - svgd.gmm_full_potential_grad(x - e, mu, covs, w, conv)[0]) / (2 * eps)
            np.testing.assert_allclose(g[:, j], fd, rtol=1e-6, atol=1e-8)
