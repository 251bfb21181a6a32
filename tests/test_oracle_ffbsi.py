"""The backward-simulation oracle (mocat/src/ssm/backward.py:20-39, 210-243) pinned by the
categorical frequencies of one backward step and, for a linear-Gaussian model, by the RTS smoother
(itself pinned by the brute-force joint-Gaussian conditional)."""
import numpy as np
import pytest

from oracle import ffbsi as of, kalman as ok, pf as opf, rng


def _model(d=2, seed=0):
    rs = np.random.default_rng(seed)
    A = 0.9 * np.eye(d) + rs.normal(0, 0.05, (d, d))
    Q, H, R = 0.3 * np.eye(d), rs.normal(0, 0.7, (d, d)), 0.5 * np.eye(d)
    return np.zeros(d), np.eye(d), A, Q, H, R


def test_rts_smoother_matches_joint_gaussian_conditional():
    m0, P0, A, Q, H, R = _model(2, 1)
    y = np.random.default_rng(2).normal(size=(5, 2))
    ms, Ps = ok.smoother(m0, P0, A, Q, H, R, y)
    mb, Pb = ok.smoother_bruteforce(m0, P0, A, Q, H, R, y)
    np.testing.assert_allclose(ms, mb, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(Ps, Pb, rtol=1e-9, atol=1e-10)


def test_one_backward_step_draws_the_backward_kernel_weights():
    """T = 2 with the final particle forced (one finite weight): over many keys the step-0 index
    frequencies equal w_j exp(-U(x_j -> x_1)) normalised (a sign error in U flips them)."""
    d = 1
    pot = of.lgssm_transition(np.array([[0.8]]), np.array([[0.5]]))
    vals = np.array([[[-1.0], [0.0], [0.4], [1.5], [2.0]], [[0.3], [9.0], [9.0], [9.0], [9.0]]])
    lw = np.array([[0.2, -0.5, 0.0, 0.7, -1.0], [0.0, -np.inf, -np.inf, -np.inf, -np.inf]])
    logits = lw[0] - pot(vals[0], vals[1][0])
    p = np.exp(logits - logits.max())
    p /= p.sum()
    K = 3000
    counts = np.zeros(5)
    for s in range(K):
        _, idx = of.backward_simulation_full(pot, vals, lw, rng.PRNGKey(s), 1)
        assert idx[1, 0] == 0
        counts[idx[0, 0]] += 1
    sd = np.sqrt(K * p * (1 - p))
    assert np.all(np.abs(counts - K * p) < 4.5 * sd + 1), (counts, K * p)
    wrong = np.exp(lw[0] + pot(vals[0], vals[1][0]))
    wrong /= wrong.sum()
    assert np.abs(counts / K - wrong).max() > 0.1


def test_backward_trajectories_converge_to_the_rts_smoother():
    m0, P0, A, Q, H, R = _model(1, 3)
    p = opf.lgssm_params(m0, P0, A, Q, H, R)
    T = 6
    _, y = opf.simulate(p, T, rng.PRNGKey(4))
    n = 400
    r = opf.run_pf(p, y, rng.PRNGKey(5), n, ess_threshold=1.0)
    traj, _ = of.backward_simulation_full(of.lgssm_transition(A, Q), r["value"], r["log_weight"], rng.PRNGKey(6), n)
    ms, Ps = ok.smoother(m0, P0, A, Q, H, R, y)
    sd = np.sqrt(Ps[:, 0, 0])
    err = np.abs(traj[:, :, 0].mean(1) - ms[:, 0]) / sd
    assert err.max() < 0.3, err
    ratio = traj[:, :, 0].std(1) / sd
    assert np.all((ratio > 0.7) & (ratio < 1.3)), ratio


def test_rejection_step_draws_the_backward_kernel_weights():
    """With the bound at the transition density's maximum (exp(-normaliser)) every accepted proposal
    is an exact draw: one step's index frequencies over many keys equal w_j exp(-U_j) normalised;
    with a single round the rest are full draws, the same distribution."""
    A, Q = np.array([[0.8]]), np.array([[0.5]])
    pot = of.lgssm_transition(A, Q)
    vals = np.array([[-1.0], [0.0], [0.4], [1.5], [2.0]])
    lw = np.array([0.2, -0.5, 0.0, 0.7, -1.0])
    x1 = np.array([[0.3]])
    logits = lw - pot(vals, x1[0])
    p = np.exp(logits - logits.max())
    p /= p.sum()
    top = float(np.exp(-(np.log(2 * np.pi) + np.log(0.5)) / 2))
    K = 2000
    for rounds in (30, 1):
        counts = np.zeros(5)
        for s in range(K):
            idx, evals = of.rejection_step(pot, vals, lw, x1, rng.PRNGKey(s), rounds, top, 1.01)
            counts[idx[0]] += 1
            assert evals >= 1
        sd = np.sqrt(K * p * (1 - p))
        assert np.all(np.abs(counts - K * p) < 4.5 * sd + 1), (rounds, counts, K * p)
