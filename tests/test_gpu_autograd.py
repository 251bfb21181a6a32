"""torch.autograd through the fused scenarios: gradients of a potential taken by torch.autograd.grad
reach the kernels' hand-written adjoints (the counterpart of jax.grad through the jax.custom_vjp the
reference's Scenario would be differentiated with, mocat/src/core.py:204-230), checked against the
adjoint itself and against central finite differences of the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import lr as olr
from oracle import svgd as osv

pytestmark = pytest.mark.gpu


def _fd(f, x, eps):
    """Central differences of a scalar FP64 function, coordinate by coordinate."""
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = eps
        out.flat[i] = (f(x + e) - f(x - e)) / (2 * eps)
    return out


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp64"])
def test_lr_autograd_reaches_the_adjoint(precision):
    import mocat_amd as m
    X, y, _ = olr.make_dataset(3, 700, 20)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision=precision)
    th0 = (0.3 * np.random.default_rng(4).standard_normal((16, 20))).astype(np.float32)
    th = torch.tensor(th0, device="cuda", requires_grad=True)
    U = sc.potential(th)
    w = torch.linspace(0.5, 2.0, 16, device="cuda")
    (g,) = torch.autograd.grad((w * U).sum(), th)
    pp, lp, gp, gl = sc.evaluate(th.detach())
    # the autograd values are the kernel's own outputs: the adjoint scaled by the incoming gradient
    assert torch.equal(U.detach(), pp + lp)
    assert torch.equal(g, w[:, None] * (gp + gl))
    # tempered potential and its parts
    Ut = sc.tempered_potential(th, 0.3)
    (gt,) = torch.autograd.grad(Ut.sum(), th)
    assert torch.equal(gt, gp + 0.3 * gl)
    (gq,) = torch.autograd.grad(sc.likelihood_potential(th).sum() + sc.prior_potential(th).sum(), th)
    assert torch.equal(gq, gl + gp)
    # against central differences of the FP64 potential (oracle), for four chains
    tol = {"fp32": 2e-5, "bf16x3": 1e-4, "fp64": 1e-6}[precision]
    gscale = np.abs(X).sum(0).max() + 1.0
    gh = g.cpu().numpy().astype(np.float64) / w.cpu().numpy()[:, None]
    for c in range(4):
        fd = _fd(lambda v: olr.potential_and_grad(v[None], X, y, 0.1)[0][0], th0[c], 1e-5)
        assert np.max(np.abs(gh[c] - fd)) / gscale < tol


def test_lr_autograd_host_tensor_and_shapes():
    import mocat_amd as m
    X, y, _ = olr.make_dataset(5, 300, 7)
    sc = m.LogisticRegression(X, y, precision="fp32")
    th = torch.zeros((2, 3, 7), requires_grad=True)  # host tensor, batch shape (2, 3)
    U = sc.potential(th)
    assert U.shape == (2, 3)
    (g,) = torch.autograd.grad(U.sum(), th)
    assert g.device == th.device and g.shape == th.shape
    want = olr.potential_and_grad(np.zeros((6, 7)), X, y, 0.1)[1].reshape(2, 3, 7)
    np.testing.assert_allclose(g.numpy(), want, rtol=1e-5, atol=1e-4)


def test_gmm_autograd_reaches_the_adjoint():
    import mocat_amd as m
    rs = np.random.default_rng(7)
    d, K = 6, 3
    mu, var, wts = rs.normal(0, 1, (K, d)), rs.uniform(0.5, 2, (K, d)), rs.dirichlet(np.ones(K))
    sc = m.GaussianMixture(mu, var, wts)
    x0 = rs.normal(0, 1.5, (10, d)).astype(np.float32)
    x = torch.tensor(x0, device="cuda", requires_grad=True)
    U = sc.potential(x)
    (g,) = torch.autograd.grad(U.sum(), x)
    _, lp, _, gl = sc.evaluate(x.detach())
    assert torch.equal(U.detach(), lp) and torch.equal(g, gl)
    for c in range(3):
        fd = _fd(lambda v: osv.gmm_potential_grad(v[None], mu, var, wts)[0][0], x0[c], 1e-5)
        np.testing.assert_allclose(g[c].cpu().numpy(), fd, rtol=1e-4, atol=1e-4)
