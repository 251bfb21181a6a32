"""GPU backward simulation (mocat/src/ssm/backward.py) vs the FP64 oracle step by step (same keys,
the GPU's own later-time trajectories as input), and the RTS smoother for a linear-Gaussian model."""
import numpy as np
import pytest
import torch

from oracle import ffbsi as of, kalman as ok, rng as orng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def _lgssm(m, d, seed):
    rs = np.random.default_rng(seed)
    A = 0.9 * np.eye(d) + rs.normal(0, 0.05, (d, d))
    Q, H, R = 0.3 * np.eye(d), rs.normal(0, 0.7, (d, d)), 0.5 * np.eye(d)
    model = m.ssm.TimeHomogenousLinearGaussian(np.zeros(d), np.eye(d), A, Q, H, R)
    return model, A, Q, H, R


# a draw whose best two scores (Gumbel + log-weight - transition potential) are within TIE of each other
# may be decided differently by the GPU's single-precision scores: every other draw must match exactly
TIE = 1e-4


def _check_steps(m, res, marg, pot, key, n_samps):
    """Every step's indices against the oracle given the GPU's trajectories at the next time: exact
    wherever the draw is not a near-tie; returns the fraction of near-ties."""
    vals = marg.value.cpu().numpy().astype(np.float64)
    lw = marg.log_weight.cpu().numpy().astype(np.float64)
    out = res.value.cpu().numpy()
    anc = res.ancestors.cpu().numpy()
    T = vals.shape[0]
    tk = orng.split(key.cpu().numpy(), T)
    want, margin = orng.categorical_margin(tk[T - 1], lw[T - 1], n_samps)
    ties = []
    for t in range(T - 1, -1, -1):
        if t < T - 1:
            want, margin = of.backward_step_margin(pot, vals[t], lw[t], out[t + 1], tk[t])
        decided = margin > TIE
        np.testing.assert_array_equal(anc[t][decided], want[decided])
        ties.append((~decided).mean())
    np.testing.assert_array_equal(out, np.take_along_axis(vals, anc[:, :, None].astype(np.int64), 1).astype(np.float32))
    return np.array(ties)


@pytest.mark.parametrize("d,n,n_samps", [(1, 1, 3), (2, 300, 200), (4, 257, 64)])
def test_lgssm_backward_steps_vs_oracle(m, d, n, n_samps):
    model, A, Q, H, R = _lgssm(m, d, d + n)
    sim = model.simulate(torch.arange(8.0), m.random.PRNGKey(1))
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), sim.y, sim.t, m.random.PRNGKey(2),
                                                   n=n, ess_threshold=0.5)
    key = m.random.PRNGKey(3)
    res = m.ssm.backward_simulation(model, marg, key, n_samps)
    assert res.value.shape == (8, n_samps, d) and not hasattr(res, "log_weight")
    np.testing.assert_array_equal(res.num_transition_evals.numpy(), np.r_[0, np.full(7, n * n_samps)])
    ties = _check_steps(m, res, marg, of.lgssm_transition(A, Q), key, n_samps)
    assert ties.max() <= 0.01, ties


def test_sv_backward_steps_vs_oracle(m):
    model = m.ssm.StochasticVolatility(transition_matrix=0.95, transition_covariance=0.3,
                                       likelihood_covariance_diag=0.8)
    sim = model.simulate(torch.arange(10.0), m.random.PRNGKey(4))
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), sim.y, sim.t, m.random.PRNGKey(5),
                                                   n=500, ess_threshold=1.0)
    key = m.random.PRNGKey(6)
    res = m.ssm.backward_simulation(model, marg, key, 300)
    ties = _check_steps(m, res, marg, of.sv_transition(0.95, np.sqrt(0.3)), key, 300)
    assert ties.max() <= 0.01, ties


def test_lgssm_trajectories_converge_to_rts_smoother(m):
    d, T = 2, 20
    model, A, Q, H, R = _lgssm(m, d, 7)
    sim = model.simulate(torch.arange(float(T)), m.random.PRNGKey(8))
    res = m.ssm.forward_filtering_backward_simulation(model, m.ssm.BootstrapFilter(), sim.y, sim.t, 4096,
                                                      m.random.PRNGKey(9), ess_threshold=0.5)
    ms, Ps = ok.smoother(np.zeros(d), np.eye(d), A, Q, H, R, sim.y.cpu().numpy())
    x = res.value.cpu().numpy().astype(np.float64)
    sd = np.sqrt(np.stack([np.diag(P) for P in Ps]))
    err = np.abs(x.mean(1) - ms) / sd
    assert err.max() < 0.2, err
    ratio = x.std(1) / sd
    assert np.all((ratio > 0.8) & (ratio < 1.2)), ratio
    assert res.time >= res.pf_time > 0 and res.bsi_time > 0


def _rejection_agreement(m, model, pot, n, n_samps, T, max_rej, bound, infl, seed):
    sim = model.simulate(torch.arange(float(T)), m.random.PRNGKey(seed))
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), sim.y, sim.t,
                                                   m.random.PRNGKey(seed + 1), n=n, ess_threshold=0.5)
    key = m.random.PRNGKey(seed + 2)
    res = m.ssm.backward_simulation(model, marg, key, n_samps, maximum_rejections=max_rej,
                                    transition_dens_bound_parameter=bound, bound_inflation=infl)
    vals = marg.value.cpu().numpy().astype(np.float64)
    lw = marg.log_weight.cpu().numpy().astype(np.float64)
    out, anc = res.value.cpu().numpy(), res.ancestors.cpu().numpy()
    tk = orng.split(key.cpu().numpy(), T)
    ties, ev = [], []
    for t in range(T - 2, -1, -1):
        idx, evals, tie = of.rejection_step(pot, vals[t], lw[t], out[t + 1], tk[t], max_rej, bound, infl, TIE)
        # exact for every trajectory none of whose draws or acceptance tests was a near-tie
        np.testing.assert_array_equal(anc[t][~tie], idx[~tie])
        ties.append(tie.mean())
        ev.append((res.num_transition_evals[t + 1].item(), evals))
    assert res.num_transition_evals[0].item() == 0
    np.testing.assert_array_equal(out, np.take_along_axis(vals, anc[:, :, None].astype(np.int64), 1).astype(np.float32))
    return np.array(ties), np.array(ev)


@pytest.mark.parametrize("max_rej,bound", [(1, 0.0), (10, 0.0), (40, 5.0)])
def test_rejection_backward_steps_vs_oracle(m, max_rej, bound):
    """Rejection rounds (backward.py:42-194): the same proposals, acceptances, bound raises and
    final full draws as the oracle on the same keys, every step; the evaluation counts follow."""
    model, A, Q, H, R = _lgssm(m, 2, 21)
    ties, ev = _rejection_agreement(m, model, of.lgssm_transition(A, Q), 300, 200, 6, max_rej, bound, 1.01, 30)
    assert ties.max() <= 0.02, ties
    assert np.all(np.abs(ev[:, 0] - ev[:, 1]) <= 0.02 * ev[:, 1] + 300), ev


def test_rejection_sv_vs_oracle(m):
    model = m.ssm.StochasticVolatility(transition_matrix=0.9, transition_covariance=0.4,
                                       likelihood_covariance_diag=0.7)
    ties, ev = _rejection_agreement(m, model, of.sv_transition(0.9, np.sqrt(0.4)), 400, 256, 6, 12, 0.0, 1.05, 40)
    assert ties.max() <= 0.02, ties


def test_rejection_trajectories_converge_to_rts_smoother(m):
    d, T = 2, 15
    model, A, Q, H, R = _lgssm(m, d, 7)
    sim = model.simulate(torch.arange(float(T)), m.random.PRNGKey(8))
    res = m.ssm.forward_filtering_backward_simulation(model, m.ssm.BootstrapFilter(), sim.y, sim.t, 2048,
                                                      m.random.PRNGKey(10), ess_threshold=0.5, maximum_rejections=20)
    ms, Ps = ok.smoother(np.zeros(d), np.eye(d), A, Q, H, R, sim.y.cpu().numpy())
    x = res.value.cpu().numpy().astype(np.float64)
    sd = np.sqrt(np.stack([np.diag(P) for P in Ps]))
    assert (np.abs(x.mean(1) - ms) / sd).max() < 0.25
    ratio = x.std(1) / sd
    assert np.all((ratio > 0.75) & (ratio < 1.25)), ratio
    # rejection needs far fewer density evaluations than the N x M of full resampling
    assert res.num_transition_evals[1:].mean().item() < 0.2 * 2048 * 2048


def test_ffbsi_key_tree(m):
    model, *_ = _lgssm(m, 2, 11)
    sim = model.simulate(torch.arange(6.0), m.random.PRNGKey(1))
    key = m.random.PRNGKey(12)
    res = m.ssm.forward_filtering_backward_simulation(model, m.ssm.BootstrapFilter(), sim.y, sim.t, 50, key, n_pf=80)
    k = m.random.split(key, 2)
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), sim.y, sim.t, k[0], n=80)
    again = m.ssm.backward_simulation(model, marg, k[1], 50)
    np.testing.assert_array_equal(res.value.cpu().numpy(), again.value.cpu().numpy())


@pytest.mark.parametrize("T,n,n_samps", [(1, 50, 70), (2, 1, 5), (3, 64, 200)])
def test_short_series_and_more_trajectories_than_particles(m, T, n, n_samps):
    """T = 1 is the final draw alone; one particle forces every index; more trajectories than
    particles draw with replacement."""
    model, A, Q, H, R = _lgssm(m, 2, T + n)
    sim = model.simulate(torch.arange(float(T)), m.random.PRNGKey(2))
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), sim.y, sim.t, m.random.PRNGKey(3),
                                                   n=n)
    key = m.random.PRNGKey(4)
    for rej in (0, 5):
        res = m.ssm.backward_simulation(model, marg, key, n_samps, maximum_rejections=rej)
        assert res.value.shape == (T, n_samps, 2)
        anc = res.ancestors.cpu().numpy()
        assert anc.min() >= 0 and anc.max() < n
        if n == 1:
            assert (anc == 0).all()
    lw = marg.log_weight.cpu().numpy().astype(np.float64)
    tk = orng.split(key.cpu().numpy(), T)
    full = m.ssm.backward_simulation(model, marg, key, n_samps)
    assert np.mean(full.ancestors.cpu().numpy()[T - 1] == orng.categorical(tk[T - 1], lw[T - 1], n_samps)) >= 0.99
