"""Forward-filtering backward-simulation, full resampling (mocat/src/ssm/backward.py:20-39, 210-243),
restated in FP64 NumPy (test infrastructure only).

  t_keys = split(key, T); x_{T-1} = particles_{T-1}[categorical(t_keys[T-1], lw_{T-1}, shape=(M,))]
  for t = T-2 .. 0, trajectory m: key_m = split(t_keys[t], M)[m],
      x_t[m] = particles_t[argmax_j(gumbel(key_m, (N,))[j] + lw_t[j] - U(particles_t[j] -> x_{t+1}[m]))]
U is mocat's transition potential through gaussian_potential (utils.py:49-105):
  linear Gaussian  0.5 |(x' - A x) @ inv(chol(Q))|^2 + (d log 2 pi + log det Q) / 2  (linear_gaussian.py:73-84)
  stochastic vol.  0.5 (x' - phi x)^2 / sigma^2 + (log 2 pi + 2 log sigma) / 2, unit steps (:78-84)
Pinned by the categorical frequencies of one backward step (the weights lw_j - U_j, a sign error
flips them) and, for a linear-Gaussian model, by the RTS smoother means the trajectories converge
to (tests/test_oracle_ffbsi.py); exact indices near f32 ties of the GPU's Gumbel-max are
"parity unpinned", as for mocat's categorical.
"""
import numpy as np

from . import rng

LOG2PI = np.log(2 * np.pi)


def lgssm_transition(A, Q):
    A = np.asarray(A, np.float64)
    Q = np.asarray(Q, np.float64)
    S = np.linalg.inv(np.linalg.cholesky(Q))
    nlz = (A.shape[0] * LOG2PI + np.log(np.linalg.det(Q))) / 2

    def pot(xp, xn):  # xp (N, d), xn (d,) -> (N,)
        diff = xn[None, :] - xp @ A.T
        return 0.5 * np.sum((diff @ S) ** 2, -1) + nlz
    return pot


def sv_transition(phi, sigma):
    nlz = (LOG2PI + 2 * np.log(sigma)) / 2

    def pot(xp, xn):
        return 0.5 * (xn[0] - phi * xp[:, 0]) ** 2 / sigma ** 2 + nlz
    return pot


def backward_simulation_full(pot, vals, log_weight, key, n_samps):
    """Returns (trajectories (T, M, d), indices (T, M))."""
    vals = np.asarray(vals, np.float64)
    lw = np.asarray(log_weight, np.float64)
    T, n, d = vals.shape
    tk = rng.split(np.asarray(key, np.uint32), T)
    idx = np.zeros((T, n_samps), np.int64)
    idx[T - 1] = rng.categorical(tk[T - 1], lw[T - 1], n_samps)
    x = vals[T - 1][idx[T - 1]]
    out = np.zeros((T, n_samps, d))
    out[T - 1] = x
    for t in range(T - 2, -1, -1):
        idx[t] = backward_step(pot, vals[t], lw[t], x, tk[t])
        x = vals[t][idx[t]]
        out[t] = x
    return out, idx


def backward_step(pot, vals_t, lw_t, x_next, key_t):
    """Indices of one backward step for the trajectories x_next (M, d) with key_t = t_keys[t]."""
    return backward_step_margin(pot, vals_t, lw_t, x_next, key_t)[0]


def backward_step_margin(pot, vals_t, lw_t, x_next, key_t):
    """backward_step plus each draw's margin between its best and second-best score (FP64)."""
    n = len(lw_t)
    keys = rng.split(np.asarray(key_t, np.uint32), len(x_next))
    out = np.zeros(len(x_next), np.int64)
    margin = np.zeros(len(x_next))
    for m in range(len(x_next)):
        logits = np.asarray(lw_t, np.float64) - pot(np.asarray(vals_t, np.float64), np.asarray(x_next[m], np.float64))
        i, mg = rng.argmax_margin((rng.gumbel(keys[m], (n,)) + logits)[None, :])
        out[m], margin[m] = int(i[0]), float(mg[0])
    return out, margin


def rejection_step(pot, vals_t, lw_t, x_next, key_t, maximum_rejections, init_bound, inflation, tie_tol=None):
    """rejection_resampling (backward.py:130-194) for one backward step: returns (indices (M,),
    transition-density evaluations). Densities, the bound and the acceptance tests in float32 as
    mocat computes them; categorical draws by Gumbel-max in FP64. With tie_tol, also returns a flag per
    trajectory: some draw of it had a score margin <= tie_tol, or some acceptance test a uniform within
    tie_tol (relative) of its threshold, so single-precision arithmetic may decide it differently."""
    res = _rejection_step(pot, vals_t, lw_t, x_next, key_t, maximum_rejections, init_bound, inflation,
                          1e-300 if tie_tol is None else tie_tol)
    return res[:2] if tie_tol is None else res


def _rejection_step(pot, vals_t, lw_t, x_next, key_t, maximum_rejections, init_bound, inflation, tie_tol):
    F = np.float32
    vals_t = np.asarray(vals_t, np.float64)
    lw_t = np.asarray(lw_t, np.float64)
    M, n = len(x_next), len(lw_t)
    k0, k1, k2 = rng.split(np.asarray(key_t, np.uint32), 3)
    dens_of = lambda i, m: F(np.exp(-pot(vals_t[[i]], np.asarray(x_next[m], np.float64))[0]))
    idx, mg = rng.categorical_margin(k0, lw_t, M)
    idx = idx.astype(np.int64)
    tie = mg <= tie_tol
    dens = np.array([dens_of(idx[m], m) for m in range(M)], F)
    mx = dens.max()
    bound = F(mx * F(inflation)) if mx > F(init_bound) else F(init_bound)
    u0 = rng.uniform(k1, (M,))
    thr0 = (dens / bound).astype(F)
    tie |= np.abs(u0.astype(np.float64) - thr0) <= tie_tol * np.maximum(thr0, 1e-30)
    wait = u0 > thr0
    keys = rng.split(k2, M)
    evals = M
    it = 1
    while wait.sum() > 0 and it < maximum_rejections:
        entered = wait.copy()
        d_round = np.zeros(M, F)
        new_wait = np.zeros(M, bool)
        for m in range(M):
            if not entered[m]:
                continue
            kr, kc, ku = rng.split(keys[m], 3)
            i, mgi = rng.argmax_margin((rng.gumbel(kc, (n,)) + lw_t)[None, :])
            idx[m] = int(i[0])
            tie[m] |= mgi[0] <= tie_tol
            d_round[m] = dens_of(idx[m], m)
            u = rng.uniform(ku, ())
            thr = F(d_round[m] / bound)
            tie[m] |= abs(float(u) - float(thr)) <= tie_tol * max(float(thr), 1e-30)
            new_wait[m] = u > thr
            keys[m] = kr
        mx = d_round.max()
        if mx > bound:
            bound = F(mx * F(inflation))
            new_wait[:] = True
        wait = new_wait
        evals += int(entered.sum())
        it += 1
    for m in range(M):
        if wait[m]:
            logits = lw_t - pot(vals_t, np.asarray(x_next[m], np.float64))
            i, mgi = rng.argmax_margin((rng.gumbel(keys[m], (n,)) + logits)[None, :])
            idx[m] = int(i[0])
            tie[m] |= mgi[0] <= tie_tol
    evals += n * int(wait.sum())
    return idx, evals, tie
