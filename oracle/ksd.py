"""Kernel Stein discrepancy (mocat/src/metrics.py:88-130) with mocat's Gaussian and IMQ kernels
(mocat/src/kernels.py:83-111, 164-215), restated pair by pair in FP64.

Test infrastructure: only tests/, the smoke check and the benchmark's CPU baseline use this module.

mocat's k_0 uses the potential gradient where the Stein kernel has the score:
  k_0(x, y) = sum(diag_grad_xy) + grad_x k . g(y) + g(x) . grad_y k + k g(x).g(y),   g = grad U,
so it equals the textbook Stein kernel evaluated with score := g. mocat's IMQ methods are restated
as written: its grad_x carries no 1/bandwidth factor and the second term of its diag_grad_xy has a
plus sign (both differ from the derivatives of its _call unless bandwidth = 1 / beta = 1).
"""
import numpy as np


def gaussian(x, y, bandwidth):
    diff = (x - y) / bandwidth
    return np.exp(-0.5 * np.sum(diff ** 2, -1))


def gaussian_grad_x(x, y, bandwidth):
    return (y - x) * gaussian(x, y, bandwidth)[..., None] / bandwidth ** 2


def gaussian_grad_y(x, y, bandwidth):
    return (x - y) * gaussian(x, y, bandwidth)[..., None] / bandwidth ** 2


def gaussian_diag_grad_xy(x, y, bandwidth):
    return (bandwidth ** 2 - (x - y) ** 2) * gaussian(x, y, bandwidth)[..., None] / bandwidth ** 4


def _imq_base(x, y, c, bandwidth):
    diff = (x - y) / bandwidth
    return diff, c + 0.5 * np.sum(diff ** 2, -1)


def imq(x, y, c, beta, bandwidth):
    _, base = _imq_base(x, y, c, bandwidth)
    return base ** beta


def imq_grad_x(x, y, c, beta, bandwidth):
    diff, base = _imq_base(x, y, c, bandwidth)
    return diff * beta * (base ** (beta - 1))[..., None]


def imq_grad_y(x, y, c, beta, bandwidth):
    diff, base = _imq_base(x, y, c, bandwidth)
    return -diff / bandwidth * beta * (base ** (beta - 1))[..., None]


def imq_diag_grad_xy(x, y, c, beta, bandwidth):
    diff, base = _imq_base(x, y, c, bandwidth)
    return (-beta * (base ** (beta - 1))[..., None]
            + beta * (beta - 1) * diff ** 2 * (base ** (beta - 2))[..., None]) / bandwidth ** 2


KERNELS = {
    "gaussian": (gaussian, gaussian_grad_x, gaussian_grad_y, gaussian_diag_grad_xy),
    "imq": (imq, imq_grad_x, imq_grad_y, imq_diag_grad_xy),
}


def k0(x, y, gx, gy, kernel="gaussian", **params):
    """mocat's k_0 for x (d,) against rows y (m, d) (metrics.py:116-124, without the weights)."""
    call, grad_x, grad_y, diag = KERNELS[kernel]
    xb = np.broadcast_to(x, y.shape)
    return (np.sum(diag(xb, y, **params), -1) + np.sum(grad_x(xb, y, **params) * gy, -1)
            + np.sum(gx * grad_y(xb, y, **params), -1) + call(xb, y, **params) * (gy @ gx))


def ksd(x, g, kernel="gaussian", log_weight=None, inds=None, **params):
    """sqrt(sum_i sum_{j in batch_i} w_i w_j k_0(x_i, x_j) / (sum(w) sum_{j in batch_i} w_j))
    (metrics.py:101-130): batch_i = all particles (ensemble_batchsize=None, inds None), or row i of
    inds (n, batch), the indices random.choice(key, n, (n, batch)) draws."""
    x = np.asarray(x, np.float64)
    g = np.asarray(g, np.float64)
    n = len(x)
    w = np.ones(n) if log_weight is None else np.exp(np.asarray(log_weight, np.float64))
    s = w.sum()
    total = 0.0
    for i in range(n):
        if inds is None:
            total += w[i] * np.sum(w * k0(x[i], x, g[i], g, kernel, **params)) / (s * s)
        else:
            b = np.asarray(inds[i], np.int64)
            total += w[i] * np.sum(w[b] * k0(x[i], x[b], g[i], g[b], kernel, **params)) / (s * w[b].sum())
    return float(np.sqrt(total))
