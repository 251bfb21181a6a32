"""Bootstrap particle filter and resampling, NumPy restatement (test infrastructure only).

Follows mocat's run_particle_filter_for_marginals (mocat/src/ssm/filtering.py:255-324) key by key:
  (key, sub) = split(key)                             :267
  initiate_particles: split(sub, n), initial_sample, -likelihood_potential     :173-193
  int_keys = split(key, T - 1)                        :278
  resample iff ESS_{t-1} < threshold * n, lw reset to 0                        :286-292
  split(int_key, n) -> transition_sample keys; lw += -likelihood_potential     :294-303
Models:
  TimeHomogenousLinearGaussian  mocat/src/ssm/linear_gaussian/linear_gaussian.py:46-51,86-94,118-128,
                                covariances factored as reset_covariance (mocat/src/utils.py:247-262)
  StochasticVolatility          mocat/src/ssm/scenarios/stochastic_volatility.py:15-108 (1-D scalar form,
                                and dim > 1 with full A, Q and a diagonal likelihood covariance)
The log-likelihood (absent in mocat) is log Z = sum_t [lse(lw_t) - lse(lw_res_{t-1})] (the SMC
log-normaliser recursion of mocat/src/transport/smc.py:212-215), t = 0 term lse(lw_0) - log n.

Resampling from given non-negative f32 weights (an addition to mocat, which only has the Gumbel-max
categorical): integer weights q_i = floor(w_i * 2^s) with s = 52 - bitlen(n) - exponent(max w)
(exact scaling), C = cumsum(q) exactly, target T_j: systematic (j + u) * (W / n), multinomial
u_j * W (FP64, IEEE), ancestor = first i with C_i > T_j. Pinned by brute-force counting on small
inputs, offspring-count bounds (systematic: floor/ceil of n w_i / W) and the FP64 ESS/lse closed
forms (tests/test_oracle_pf.py). The PF log-likelihood is pinned against the exact Kalman
log-likelihood (oracle.kalman) statistically.
"""
import numpy as np

from . import rng

LOG2PI = np.log(2 * np.pi)


def lgssm_params(initial_mean, initial_covariance, transition_matrix, transition_covariance,
                 likelihood_matrix, likelihood_covariance):
    """Kernel parameter layout [m0, L0, A, LQ, H, SR, nlz] with mocat's covariance factorisation."""
    m0 = np.asarray(initial_mean, np.float64)
    P0 = np.asarray(initial_covariance, np.float64)
    A = np.asarray(transition_matrix, np.float64)
    Q = np.asarray(transition_covariance, np.float64)
    H = np.asarray(likelihood_matrix, np.float64)
    R = np.asarray(likelihood_covariance, np.float64)
    L0 = np.linalg.cholesky(P0)
    LQ = np.linalg.cholesky(Q)
    LR = np.linalg.cholesky(R)
    SR = np.linalg.inv(LR)                 # likelihood_precision_sqrt
    det_prec = 1.0 / np.linalg.det(R)      # likelihood_precision_det
    dy = H.shape[0]
    nlz = (dy * LOG2PI - np.log(det_prec)) / 2
    return dict(m0=m0, L0=L0, A=A, LQ=LQ, H=H, SR=SR, nlz=nlz, d=m0.shape[0], dy=dy)


def lgssm_flat(p):
    return np.concatenate([p["m0"].ravel(), p["L0"].ravel(), p["A"].ravel(), p["LQ"].ravel(), p["H"].ravel(),
                           p["SR"].ravel(), [p["nlz"]]]).astype(np.float32)


def sv_params(phi, sigma, beta, m0=0.0, var0=1.0):
    return dict(m0=np.array([m0]), L0=np.array([[np.sqrt(var0)]]), phi=float(phi), sigma=float(sigma),
                inv_beta=1.0 / float(beta), d=1, dy=1, sv=True)


def svd_params(m0, P0, A, Q, rdiag):
    """Multivariate stochastic volatility (mocat/src/ssm/scenarios/stochastic_volatility.py:15-108):
    x0 = L0 eps + m0, x' = A x + (LQ eps) s, lik = gaussian_potential(y, 0, sqrt_prec = r^-1/2 / exp(x/2))
    with L0 = chol(P0), LQ = chol(Q) (reset_covariance, mocat/src/utils.py:247-262)."""
    m0 = np.asarray(m0, np.float64).reshape(-1)
    d = m0.size
    return dict(m0=m0, L0=np.linalg.cholesky(np.asarray(P0, np.float64)), A=np.asarray(A, np.float64),
                LQ=np.linalg.cholesky(np.asarray(Q, np.float64)),
                ib=1.0 / np.sqrt(np.asarray(rdiag, np.float64).reshape(d)), d=d, dy=d, svd=True)


def svd_flat(p):
    return np.concatenate([p["m0"], p["L0"].ravel(), p["A"].ravel(), p["LQ"].ravel(), p["ib"]]).astype(np.float32)


def sv_flat(p):
    return np.array([p["m0"][0], p["L0"][0, 0], p["phi"], p["sigma"], p["inv_beta"]], np.float32)


def init_state(p, eps):
    if p.get("sv"):
        return eps * p["L0"][0, 0] + p["m0"][0]
    return eps @ p["L0"].T + p["m0"]


def transition(p, x, eps, scale=1.0):
    """One step; for stochastic volatility the noise scales with scale = sqrt(t_new - t_previous)
    (mocat/src/ssm/scenarios/stochastic_volatility.py:86-93: normal @ cov_sqrt.T * sqrt(dt)); the
    linear-Gaussian model is time homogeneous."""
    if p.get("sv"):
        return x * p["phi"] + (eps * p["sigma"]) * scale
    if p.get("svd"):
        return x @ p["A"].T + (eps @ p["LQ"].T) * scale
    return x @ p["A"].T + eps @ p["LQ"].T


def lik_potential(p, x, y):
    if p.get("sv"):
        s = p["inv_beta"] / np.exp(0.5 * x[:, 0])
        return 0.5 * (y[0] * s) ** 2 + 0.5 * LOG2PI - np.log(s)
    if p.get("svd"):
        s = p["ib"][None, :] / np.exp(0.5 * x)
        return np.sum(0.5 * (y[None, :] * s) ** 2 + 0.5 * LOG2PI - np.log(s), axis=1)
    diff = y[None, :] - x @ p["H"].T
    v = diff @ p["SR"]
    return 0.5 * np.sum(v * v, axis=1) + p["nlz"]


def lse(w):
    w = np.asarray(w, np.float64)
    m = np.max(w)
    if not np.isfinite(m):
        return m
    return m + np.log(np.sum(np.exp(w - m)))


def log_ess(w):
    w = np.asarray(w, np.float64)
    return 2 * lse(w) - lse(2 * w)


# ----------------------------------------------------------------------------- resampling
def quantise(w, n=None):
    w = np.asarray(w, np.float32)
    n = w.shape[0] if n is None else n
    wmax = np.float32(np.max(np.where(w > 0, w, 0)))
    _, e = np.frexp(wmax)
    shift = (52 - int(n).bit_length()) - int(e)
    wd = np.where(w > 0, w, 0).astype(np.float64)
    return np.floor(np.ldexp(wd, shift)).astype(np.uint64)


def _search(C, T):
    idx = np.searchsorted(C.astype(np.float64), T, side="right")
    return np.minimum(idx, C.shape[0] - 1).astype(np.int32)


def systematic(w, u):
    q = quantise(w)
    C = np.cumsum(q, dtype=np.uint64)
    n = q.shape[0]
    W = float(C[-1])
    step = W / n
    T = (np.arange(n, dtype=np.float64) + np.float64(u)) * step
    return _search(C, T)


def multinomial(w, us):
    q = quantise(w)
    C = np.cumsum(q, dtype=np.uint64)
    W = float(C[-1])
    T = np.asarray(us, np.float32).astype(np.float64) * W
    return _search(C, T)


def resample_indices(w, key, method):
    """Ancestors for weights w with the uniforms the kernels draw from `key`."""
    n = np.asarray(w).shape[0]
    if method == "systematic":
        return systematic(w, rng.uniform(key))
    if method == "multinomial":
        return multinomial(w, rng.uniform(key, (n,)))
    raise ValueError(method)


# ----------------------------------------------------------------------------- filter
def run_pf(p, y, key, n, ess_threshold=0.5, resampling="systematic", dtype=np.float64, ancestors=None,
           resample_flags=None, initial=None, t=None):
    """Returns dict(value (T, n, d), log_weight (T, n), ess (T,), log_likelihood, ancestors (T, n)).

    ancestors / resample_flags (optional, (T, n) and (T,)): replay given resampling decisions and
    indices instead of drawing them (the fixed-resampling-index path of another run).
    initial = (x (n, d), log_weight (n,)): resume from a sample as run_particle_filter_for_marginals
    with initial_sample (mocat/src/ssm/filtering.py:264-278): row 0 of y is not read, the step keys
    are split(key, T - 1) (no (key, sub) split) and the log-likelihood starts at 0.
    t (optional, (T,)): observation times; the stochastic-volatility step into row t has noise scale
    sqrt(t[t] - t[t-1]) (unit steps without t)."""
    y = np.asarray(y, np.float64).reshape(len(y), -1)
    T = y.shape[0]
    d = p["d"]
    key = np.asarray(key, np.uint32)
    if initial is not None:
        rkey = key
        x = np.asarray(initial[0], np.float64).reshape(n, d)
        lw = np.asarray(initial[1], np.float64).reshape(n)
    else:
        rkey, sub = rng.split(key, 2)
        eps = rng.normal_batch(rng.split(sub, n), d).astype(dtype)
        x = init_state(p, eps).reshape(n, d)
        lw = -lik_potential(p, x, y[0])
    values, lws, esss, ancs = [x], [lw], [], [np.arange(n, dtype=np.int32)]
    l0, l2 = lse(lw), lse(2 * lw)
    esss.append(np.exp(2 * l0 - l2))
    loglik = 0.0 if initial is not None else l0 - np.log(n)
    int_keys = rng.split(rkey, T - 1) if T > 1 else np.zeros((0, 2), np.uint32)
    for tt in range(1, T):
        ik = int_keys[tt - 1]
        res = esss[-1] < ess_threshold * n if resample_flags is None else bool(resample_flags[tt])
        if res:
            if ancestors is not None:
                a = np.asarray(ancestors[tt], np.int64)
            else:
                w = np.exp(np.asarray(lw, np.float64) - l0).astype(np.float32)
                a = resample_indices(w, ik, resampling)
            lw_res = np.zeros(n)
            lres = np.log(n)
        else:
            a = np.arange(n, dtype=np.int32)
            lw_res = lw
            lres = l0
        eps = rng.normal_batch(rng.split(ik, n), d).astype(dtype)
        scale = 1.0 if t is None else float(np.sqrt(float(t[tt]) - float(t[tt - 1])))
        x = transition(p, x[a], eps, scale).reshape(n, d)
        lw = lw_res - lik_potential(p, x, y[tt])
        l0, l2 = lse(lw), lse(2 * lw)
        loglik += l0 - lres
        esss.append(np.exp(2 * l0 - l2))
        values.append(x)
        lws.append(lw)
        ancs.append(a)
    return dict(value=np.stack(values), log_weight=np.stack(lws), ess=np.array(esss), log_likelihood=loglik,
                ancestors=np.stack(ancs))


def simulate(p, T, key):
    """StateSpaceModel.simulate (mocat/src/ssm/ssm.py:138-163): keys split(key, 2T), latent from
    the first T keys (initial_sample then transition_sample), observations likelihood_sample with the
    last T keys."""
    keys = rng.split(np.asarray(key, np.uint32), 2 * T)
    d, dy = p["d"], p["dy"]
    xs = []
    x = init_state(p, rng.normal(keys[0], (d,)).astype(np.float64)[None, :]).reshape(1, d)
    xs.append(x[0])
    for t in range(1, T):
        x = transition(p, x, rng.normal(keys[t], (d,)).astype(np.float64)[None, :]).reshape(1, d)
        xs.append(x[0])
    xs = np.stack(xs)
    ys = []
    for t in range(T):
        e = rng.normal(keys[T + t], (dy,)).astype(np.float64)
        if p.get("sv"):
            ys.append((1.0 / p["inv_beta"]) * np.exp(0.5 * xs[t]) * e)
        elif p.get("svd"):
            ys.append(np.exp(0.5 * xs[t]) * e / p["ib"])
        else:
            LR = np.linalg.inv(p["SR"])
            ys.append(p["H"] @ xs[t] + LR @ e)
    return xs, np.stack(ys)


def sv_loglik_grad(phi, sigma, beta, y, key, n, ancestors, resample_flags, m0=0.0, var0=1.0, t=None):
    """d log Z / d(phi, sigma, beta) of the fixed-resampling-index estimator (ancestors and noise held
    fixed) for the 1-D stochastic-volatility filter, by forward tangents in FP64:
      x_t = phi x_{a} + sigma s eps,  dx/dphi = x_a + phi dx_a/dphi,  dx/dsigma = s eps + phi dx_a/dsigma
      (s = sqrt(t_t - t_{t-1}) with observation times t, else 1)
      lik = 0.5 (y s)^2 + 0.5 log 2pi - log s, s = exp(-x/2) / beta:
          dlik/dx = 0.5 - 0.5 (y s)^2,  dlik/dbeta = (1 - (y s)^2) / beta
      d log Z = sum_t [avg_{w_t}(dlw_t) - avg_{w_res}(dlw_res)]
    Pinned by central finite differences of run_pf on the same fixed path (tests/test_oracle_pf.py)."""
    y = np.asarray(y, np.float64).reshape(-1)
    T = y.shape[0]
    rkey, sub = rng.split(np.asarray(key, np.uint32), 2)
    eps = rng.normal_batch(rng.split(sub, n), 1).astype(np.float64)[:, 0]
    x = eps * np.sqrt(var0) + m0
    dxp = np.zeros(n)
    dxs = np.zeros(n)

    def lik_parts(x, yt):
        s2 = (yt * yt) * np.exp(-x) / (beta * beta)
        lik = 0.5 * s2 + 0.5 * LOG2PI + np.log(beta) + 0.5 * x
        return lik, 0.5 - 0.5 * s2, (1.0 - s2) / beta

    lik, dldx, dldb = lik_parts(x, y[0])
    lw = -lik
    dlw = np.stack([-dldx * dxp, -dldx * dxs, -dldb], 1)
    w = np.exp(lw - lse(lw))
    grad = w @ dlw
    int_keys = rng.split(rkey, T - 1) if T > 1 else np.zeros((0, 2), np.uint32)
    for tt in range(1, T):
        ik = int_keys[tt - 1]
        s = 1.0 if t is None else float(np.sqrt(float(t[tt]) - float(t[tt - 1])))
        if resample_flags[tt]:
            a = np.asarray(ancestors[tt], np.int64)
            lw_res, dlw_res = np.zeros(n), np.zeros((n, 3))
            res_term = np.zeros(3)
        else:
            a = np.arange(n)
            lw_res, dlw_res = lw, dlw
            res_term = np.exp(lw - lse(lw)) @ dlw
        e = rng.normal_batch(rng.split(ik, n), 1).astype(np.float64)[:, 0]
        xa, dpa, dsa = x[a], dxp[a], dxs[a]
        x = phi * xa + (sigma * e) * s
        dxp = xa + phi * dpa
        dxs = e * s + phi * dsa
        lik, dldx, dldb = lik_parts(x, y[tt])
        lw = lw_res - lik
        dlw = dlw_res - np.stack([dldx * dxp, dldx * dxs, dldb], 1)
        w = np.exp(lw - lse(lw))
        grad = grad + w @ dlw - res_term
    return grad
