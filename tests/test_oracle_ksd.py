"""The KSD oracle (mocat/src/metrics.py:88-130) pinned by mocat's kernel KATs
(mocat/src/tests/test_kernels.py:16-45), finite differences of the kernel functions, the one-particle
closed form and Stein's identity under Gauss-Hermite quadrature."""
import numpy as np
import pytest

from oracle import ksd as ok


def test_kernel_kats():
    z, o = np.zeros(5), np.ones(5)
    assert ok.gaussian(z, z, 1.0) == pytest.approx(1.0)
    assert ok.gaussian(z, o, 1.0) == pytest.approx(0.082085006, abs=1e-6)
    np.testing.assert_allclose(ok.gaussian_grad_x(z, o, 1.0), o * 0.082085006, atol=1e-6)
    np.testing.assert_allclose(ok.gaussian_grad_y(z, o, 1.0), o * -0.082085006, atol=1e-6)
    assert ok.imq(z, z, 1.0, -0.5, 1.0) == pytest.approx(1.0)
    assert ok.imq(z, o, 1.0, -0.5, 1.0) == pytest.approx(0.5345225, abs=1e-6)
    np.testing.assert_allclose(ok.imq_grad_x(z, o, 1.0, -0.5, 1.0), o * 0.07636035, atol=1e-6)
    np.testing.assert_allclose(ok.imq_grad_y(z, o, 1.0, -0.5, 1.0), o * -0.07636035, atol=1e-6)


def _fd_grad(f, x, eps=1e-6):
    return np.array([(f(x + eps * e) - f(x - eps * e)) / (2 * eps) for e in np.eye(len(x))])


def _fd_diag(f, x, y, eps=1e-4):
    # d^2 f / dx_k dy_k for each k
    out = []
    for e in np.eye(len(x)):
        out.append((f(x + eps * e, y + eps * e) - f(x + eps * e, y - eps * e) - f(x - eps * e, y + eps * e)
                    + f(x - eps * e, y - eps * e)) / (4 * eps * eps))
    return np.array(out)


def test_gaussian_derivatives_finite_differences():
    rs = np.random.default_rng(0)
    x, y, h = rs.normal(size=4), rs.normal(size=4), 1.7
    np.testing.assert_allclose(ok.gaussian_grad_x(x, y, h), _fd_grad(lambda v: ok.gaussian(v, y, h), x), rtol=1e-6)
    np.testing.assert_allclose(ok.gaussian_grad_y(x, y, h), _fd_grad(lambda v: ok.gaussian(x, v, h), y), rtol=1e-6)
    np.testing.assert_allclose(ok.gaussian_diag_grad_xy(x, y, h), _fd_diag(lambda a, b: ok.gaussian(a, b, h), x, y),
                               rtol=1e-5, atol=1e-8)


def test_imq_derivatives_against_written_source():
    """grad_y is the derivative of _call; grad_x is bandwidth times it and the second term of
    diag_grad_xy has the opposite sign (both as written in mocat/src/kernels.py:184-215)."""
    rs = np.random.default_rng(1)
    x, y, c, beta, h = rs.normal(size=3), rs.normal(size=3), 0.7, -0.6, 1.9
    f = lambda a, b: ok.imq(a, b, c, beta, h)
    np.testing.assert_allclose(ok.imq_grad_y(x, y, c, beta, h), _fd_grad(lambda v: f(x, v), y), rtol=1e-6)
    np.testing.assert_allclose(ok.imq_grad_x(x, y, c, beta, h), h * _fd_grad(lambda v: f(v, y), x), rtol=1e-6)
    diff = (x - y) / h
    base = c + 0.5 * np.sum(diff ** 2)
    quirk = 2 * beta * (beta - 1) * diff ** 2 * base ** (beta - 2) / h ** 2
    np.testing.assert_allclose(ok.imq_diag_grad_xy(x, y, c, beta, h), _fd_diag(f, x, y) + quirk, rtol=1e-5,
                               atol=1e-8)


def test_single_particle_closed_form():
    rs = np.random.default_rng(2)
    x, g = rs.normal(size=(1, 6)), rs.normal(size=(1, 6))
    h = 1.3
    assert ok.ksd(x, g, "gaussian", bandwidth=h) == pytest.approx(np.sqrt(6 / h ** 2 + np.sum(g ** 2)), rel=1e-12)
    c, beta = 0.8, -0.5
    want = np.sqrt(-6 * beta * c ** (beta - 1) / h ** 2 + c ** beta * np.sum(g ** 2))
    assert ok.ksd(x, g, "imq", c=c, beta=beta, bandwidth=h) == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("h", [0.7, 1.0, 2.5])
def test_gaussian_k0_stein_identity(h):
    """With the score of N(0, I) passed as the gradient, mocat's k_0 is the Stein kernel of N(0, I):
    its expectation over one argument vanishes (2-D tensor Gauss-Hermite quadrature)."""
    nodes, wts = np.polynomial.hermite_e.hermegauss(60)
    wts = wts / wts.sum()
    X = np.stack(np.meshgrid(nodes, nodes, indexing="ij"), -1).reshape(-1, 2)
    W = np.outer(wts, wts).reshape(-1)
    for y in (np.array([0.3, -1.2]), np.array([2.0, 0.5])):
        e = np.sum(W * ok.k0(y, X, -y, -X, "gaussian", bandwidth=h))
        assert abs(e) < 1e-10, e
        # the sign a mistake would flip: with grad U (= +x) in place of the score the mean is nonzero
        assert abs(np.sum(W * ok.k0(y, X, y, X, "gaussian", bandwidth=h))) > 1e-3


def test_ksd_decreases_with_ensemble_size():
    """mocat/src/tests/test_metrics.py:88-115 pattern (dim 2, 10 vs 1000 draws, grad U = x). With
    grad U in the score's place mocat's statistic tends to a positive constant, so the comparison
    rests on the small-ensemble bias; with the score (-x) it falls like n^-1/2."""
    rs = np.random.default_rng(3)
    small, large = rs.normal(size=(10, 2)), rs.normal(size=(1000, 2))
    for kern, prm in (("gaussian", dict(bandwidth=1.0)), ("imq", dict(c=1.0, beta=-0.5, bandwidth=1.0)),
                      ("gaussian", dict(bandwidth=10.0)), ("imq", dict(c=0.1, beta=-0.1, bandwidth=10.0))):
        assert ok.ksd(large, large, kern, **prm) < ok.ksd(small, small, kern, **prm)
    k_small, k_large = ok.ksd(small, -small, "gaussian", bandwidth=1.0), ok.ksd(large, -large, "gaussian",
                                                                                  bandwidth=1.0)
    assert k_large < 0.1 and k_large < 0.5 * k_small


def test_minibatch_with_every_particle_in_each_batch_is_the_full_ksd():
    """metrics.py:101-128: batches that hold every particle once reduce to the full-ensemble sum
    (the batch weight sum becomes sum(w)); a repeated index counts twice."""
    rs = np.random.default_rng(5)
    x, g = rs.normal(size=(12, 3)), rs.normal(size=(12, 3))
    lw = rs.normal(size=12)
    inds = np.tile(np.arange(12), (12, 1))
    for kern, prm in (("gaussian", dict(bandwidth=1.3)), ("imq", dict(c=0.7, beta=-0.4, bandwidth=1.1))):
        assert ok.ksd(x, g, kern, log_weight=lw, inds=inds, **prm) == pytest.approx(
            ok.ksd(x, g, kern, log_weight=lw, **prm), rel=1e-12)
    # a batch of one particle j: w_i k_0(x_i, x_j) / sum(w) summed over i
    j = np.full((12, 1), 4)
    w = np.exp(lw)
    want = np.sqrt(sum(w[i] * ok.k0(x[i], x[[4]], g[i], g[[4]], "gaussian", bandwidth=1.3)[0] for i in range(12))
                   / w.sum())
    assert ok.ksd(x, g, "gaussian", log_weight=lw, inds=j, bandwidth=1.3) == pytest.approx(want, rel=1e-12)
