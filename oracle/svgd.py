"""SVGD pieces restated in FP64 NumPy (test infrastructure only).

phi_pairwise: mocat's kernelised_grad_matrix (mocat/src/transport/svgd.py:18-32) summed pair by pair
  phi_i = mean_j [ -k(x_j, x_i) gradU_j + (x_i - x_j) k(x_j, x_i) / h^2 ],
  with the Gaussian kernel of mocat/src/kernels.py:89-102.
adagrad: jax.example_libraries.optimizers.adagrad(step, momentum=0.9) restated (third-party, not in
  the reference tree; called at mocat/src/transport/svgd.py:56,104-106,138-139) from its published
  update rule g_sq += g^2; m = (1 - mom) g / sqrt(g_sq) + mom m; x -= step m, pinned by closed forms
  (first step, invariance to gradient scale, constant-gradient momentum sum; tests/test_oracle_svgd.py).
gmm: mocat's GaussianMixture (mocat/src/scenarios/toy_examples.py:54-116) with diagonal covariances,
  in log-sum-exp form; det_convention "mocat" weights components by det(P), "textbook" by sqrt(det P).
gmm_full: the same with full covariances. mocat's component potential (toy_examples.py:95-100) is
  0.5 |(x - mu) @ L^T|^2 - log(w det(P) / (2 pi)^(d/2)) with L = cholesky(inv(Sigma)) lower, i.e. the
  quadratic form of L^T L (which equals P only when the factors commute, e.g. diagonal Sigma);
  "mocat" keeps that form and det(P), "textbook" uses P and sqrt(det P) (a normalised mixture).
median / mean bandwidth: mocat/src/kernels.py:220-229 with numpy's median of the full distance matrix.
Pinned by mocat's kernel and GaussianMixture known-answer tests (mocat/src/tests/test_kernels.py:16-28,
mocat/src/tests/test_toy_scenarios.py:59-106) and a brute-force check of the GEMM form of phi.
"""
import numpy as np


def gaussian_kernel(x, y, h):
    diff = (np.asarray(x, np.float64) - np.asarray(y, np.float64)) / h
    return np.exp(-0.5 * np.sum(diff * diff, axis=-1))


def phi_pairwise(x, g, h):
    x = np.asarray(x, np.float64)
    g = np.asarray(g, np.float64)
    n = x.shape[0]
    out = np.zeros_like(x)
    for i in range(n):
        k = gaussian_kernel(x, x[i][None, :], h)            # k(x_j, x_i) for all j
        out[i] = np.mean(-k[:, None] * g + (x[i][None, :] - x) * k[:, None] / h ** 2, axis=0)
    return out


def phi_gemm(x, g, h):
    x = np.asarray(x, np.float64)
    g = np.asarray(g, np.float64)
    n = x.shape[0]
    sq = np.sum(x * x, 1)
    D = np.maximum(sq[:, None] + sq[None, :] - 2 * x @ x.T, 0)
    K = np.exp(-D / (2 * h * h))
    return (-(K @ g) + (K.sum(1)[:, None] * x - K @ x) / h ** 2) / n


def adagrad(x, g_sq, m, grad, step, momentum=0.9):
    g_sq = g_sq + grad * grad
    inv = np.where(g_sq > 0, 1.0 / np.sqrt(np.where(g_sq > 0, g_sq, 1.0)), 0.0)
    m = (1 - momentum) * (grad * inv) + momentum * m
    return x - step * m, g_sq, m


def gmm_logc(weights, var, det_convention="textbook"):
    w = np.asarray(weights, np.float64)
    lam = 1.0 / np.asarray(var, np.float64)
    d = lam.shape[1]
    f = {"textbook": 0.5, "mocat": 1.0}[det_convention]
    return np.log(w) + f * np.log(lam).sum(1) - 0.5 * d * np.log(2 * np.pi)


def gmm_potential_grad(x, means, var, weights, det_convention="textbook"):
    x = np.asarray(x, np.float64)
    mu = np.asarray(means, np.float64)
    lam = 1.0 / np.asarray(var, np.float64)
    logc = gmm_logc(weights, var, det_convention)
    q = logc[None, :] - 0.5 * np.sum(lam[None] * (x[:, None, :] - mu[None]) ** 2, axis=2)
    m = q.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(q - m).sum(1))
    r = np.exp(q - lse[:, None])
    grad = np.sum(r[:, :, None] * lam[None] * (x[:, None, :] - mu[None]), axis=1)
    return -lse, grad


def gmm_full_matrices(covs, weights, det_convention="textbook"):
    """(S_k, logc_k): component k's potential is 0.5 r^T S_k r - logc_k with r = x - mu_k."""
    covs = np.asarray(covs, np.float64)
    K, d, _ = covs.shape
    P = np.linalg.inv(covs)
    L = np.linalg.cholesky(0.5 * (P + np.transpose(P, (0, 2, 1))))
    detP = np.linalg.det(P)
    w = np.asarray(weights, np.float64)
    if det_convention == "mocat":
        S = np.einsum("kji,kjl->kil", L, L)          # L^T L
        logc = np.log(w * detP) - 0.5 * d * np.log(2 * np.pi)
    else:
        S = P
        logc = np.log(w) + 0.5 * np.log(detP) - 0.5 * d * np.log(2 * np.pi)
    return S, logc


def gmm_full_potential_grad(x, means, covs, weights, det_convention="textbook"):
    x = np.asarray(x, np.float64)
    mu = np.asarray(means, np.float64)
    S, logc = gmm_full_matrices(covs, weights, det_convention)
    r = x[:, None, :] - mu[None]                         # [n][K][d]
    Sr = np.einsum("kij,nkj->nki", S, r)
    q = logc[None, :] - 0.5 * np.sum(r * Sr, axis=2)
    m = q.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(q - m).sum(1))
    wk = np.exp(q - lse[:, None])
    return -lse, np.sum(wk[:, :, None] * Sr, axis=1)


def distance_matrix(x):
    x = np.asarray(x, np.float64)
    return np.sqrt(np.maximum(np.sum((x[:, None, :] - x[None, :, :]) ** 2, -1), 0))


def median_bandwidth(x):
    D = distance_matrix(x)
    return np.median(D) / np.sqrt(2 * np.log(D.shape[0]))


def mean_bandwidth(x):
    D = distance_matrix(x)
    return np.mean(D) / np.sqrt(2 * np.log(D.shape[0]))
