"""Exact Kalman filter log-likelihood in FP64 (test infrastructure only).

mocat's run_kalman_filter_for_marginals (mocat/src/ssm/linear_gaussian/kalman.py:16-57) returns
means and covariances only, and uses the initial covariance square root as the covariance at t = 0
(kalman.py:20; exact only for P0 = I). This restatement uses the true P0 and adds the
prediction-error decomposition log p(y_0:T) = sum_t log N(y_t; H m_t|t-1, H P_t|t-1 H^T + R).
Pinned by the brute-force joint Gaussian log-density of y_0:T for small T (tests/test_oracle_pf.py).
"""
import numpy as np


def loglik(m0, P0, A, Q, H, R, y):
    m = np.asarray(m0, np.float64)
    P = np.asarray(P0, np.float64)
    A, Q, H, R = (np.asarray(v, np.float64) for v in (A, Q, H, R))
    y = np.asarray(y, np.float64)
    ll = 0.0
    means, covs = [], []
    for t in range(y.shape[0]):
        if t > 0:
            m = A @ m
            P = A @ P @ A.T + Q
        S = H @ P @ H.T + R
        r = y[t] - H @ m
        Sinv = np.linalg.inv(S)
        _, logdet = np.linalg.slogdet(S)
        ll += -0.5 * (r @ Sinv @ r + logdet + len(r) * np.log(2 * np.pi))
        K = P @ H.T @ Sinv
        m = m + K @ r
        P = P - K @ H @ P
        means.append(m.copy())
        covs.append(P.copy())
    return ll, np.stack(means), np.stack(covs)


def joint_loglik_bruteforce(m0, P0, A, Q, H, R, y):
    """log N(vec(y); mean, Cov) of the stacked observations, built from the state-space recursion."""
    A, Q, H, R = (np.asarray(v, np.float64) for v in (A, Q, H, R))
    T, dy = y.shape
    d = len(m0)
    means_x = [np.asarray(m0, np.float64)]
    for t in range(1, T):
        means_x.append(A @ means_x[-1])
    # Cov(x_s, x_t) for s <= t = A^{t-s} Cov(x_s)
    covx = [np.asarray(P0, np.float64)]
    for t in range(1, T):
        covx.append(A @ covx[-1] @ A.T + Q)
    big = np.zeros((T * dy, T * dy))
    for s in range(T):
        for t in range(s, T):
            cst = np.linalg.matrix_power(A, t - s) @ covx[s]
            blk = H @ cst.T @ H.T if True else None
            big[t * dy:(t + 1) * dy, s * dy:(s + 1) * dy] = H @ cst @ H.T
            big[s * dy:(s + 1) * dy, t * dy:(t + 1) * dy] = (H @ cst @ H.T).T
        big[s * dy:(s + 1) * dy, s * dy:(s + 1) * dy] += R
    mu = np.concatenate([H @ mx for mx in means_x])
    r = y.ravel() - mu
    _, logdet = np.linalg.slogdet(big)
    return -0.5 * (r @ np.linalg.solve(big, r) + logdet + len(r) * np.log(2 * np.pi))


def smoother(m0, P0, A, Q, H, R, y):
    """Rauch-Tung-Striebel smoothed means and covariances of x_t given y_0:T (FP64): the target
    that backward simulation samples, in the limit of many particles."""
    A, Q = (np.asarray(v, np.float64) for v in (A, Q))
    _, mf, Pf = loglik(m0, P0, A, Q, H, R, y)
    T = len(mf)
    ms, Ps = mf.copy(), Pf.copy()
    for t in range(T - 2, -1, -1):
        Pp = A @ Pf[t] @ A.T + Q
        G = Pf[t] @ A.T @ np.linalg.inv(Pp)
        ms[t] = mf[t] + G @ (ms[t + 1] - A @ mf[t])
        Ps[t] = Pf[t] + G @ (Ps[t + 1] - Pp) @ G.T
    return ms, Ps


def smoother_bruteforce(m0, P0, A, Q, H, R, y):
    """E[x_t | y_0:T] and Cov from the joint Gaussian of (x_0:T, y_0:T) conditioned on y (small T)."""
    A, Q, H, R = (np.asarray(v, np.float64) for v in (A, Q, H, R))
    y = np.asarray(y, np.float64)
    T, dy = y.shape
    d = len(m0)
    mx = [np.asarray(m0, np.float64)]
    cx = [np.asarray(P0, np.float64)]
    for t in range(1, T):
        mx.append(A @ mx[-1])
        cx.append(A @ cx[-1] @ A.T + Q)
    Sxx = np.zeros((T * d, T * d))
    for s in range(T):
        for t in range(s, T):
            c = np.linalg.matrix_power(A, t - s) @ cx[s]  # Cov(x_t, x_s)
            Sxx[t * d:(t + 1) * d, s * d:(s + 1) * d] = c
            Sxx[s * d:(s + 1) * d, t * d:(t + 1) * d] = c.T
    Hb = np.kron(np.eye(T), H)
    Syy = Hb @ Sxx @ Hb.T + np.kron(np.eye(T), R)
    Sxy = Sxx @ Hb.T
    mux = np.concatenate(mx)
    K = Sxy @ np.linalg.inv(Syy)
    mean = mux + K @ (y.ravel() - Hb @ mux)
    cov = Sxx - K @ Sxy.T
    return mean.reshape(T, d), np.stack([cov[t * d:(t + 1) * d, t * d:(t + 1) * d] for t in range(T)])
