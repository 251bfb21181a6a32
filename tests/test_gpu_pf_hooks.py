"""The batched particle-filter hooks (mocat/src/ssm/filtering.py:65-78, 124-139) and the per-particle
StateSpaceModel functions (mocat/src/ssm/ssm.py:47-81) on the compiled models: a Python time loop of
the hooks reproduces the fused filter bit for bit (same keys, no resampling), the likelihood potential
matches the oracle, and anything without a kernel raises the documented error."""
import numpy as np
import pytest
import torch

from oracle import pf as opf
from oracle import rng as orng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def m():
    import mocat_amd
    return mocat_amd


def _hook_loop(m, model, pf, y, key, n):
    """mocat's run_particle_filter_for_marginals key tree (filtering.py:264-303) without resampling,
    each step through the batched hooks."""
    T = y.shape[0]
    k, sub = m.random.split(key, 2)
    x, lw = pf.initial_sample_and_weight_vectorised(model, y[0], 0.0, m.random.split(sub, n))
    if isinstance(pf, m.ssm.OptimalNonLinearGaussianParticleFilter):
        lw = torch.zeros_like(lw)
    xs, lws = [x], [lw]
    int_keys = m.random.split(k, T - 1)
    for t in range(1, T):
        x, inc = pf.propose_and_intermediate_weight_vectorised(model, x, float(t - 1), y[t], float(t),
                                                               m.random.split(int_keys[t - 1], n))
        lw = lw + inc
        xs.append(x)
        lws.append(lw)
    return torch.stack(xs), torch.stack(lws)


@pytest.mark.parametrize("case", ["lgssm", "sv", "optimal"])
def test_hook_loop_repeats_the_fused_filter(m, case):
    from mocat_amd import datasets
    n = 777
    if case == "sv":
        models, ys, _ = datasets.sv_batch_config(2, n_models=1, t_min=25, t_max=25)
        model, y = models[0], torch.tensor(ys[0], device="cuda")
        t = torch.arange(y.shape[0], dtype=torch.float32)
    else:
        model, y, t = datasets.lgssm_config(0, 4, 25)
    pf = m.ssm.OptimalNonLinearGaussianParticleFilter() if case == "optimal" else m.ssm.BootstrapFilter()
    key = m.random.PRNGKey(9)
    out = m.ssm.run_particle_filter_for_marginals(model, pf, y, t, key, n=n, ess_threshold=0.0,
                                                  resampling="systematic")
    xs, lws = _hook_loop(m, model, pf, y, key, n)
    assert torch.equal(xs, out.value) and torch.equal(lws, out.log_weight)


def test_state_space_model_functions(m):
    from mocat_amd import datasets
    model, y, _ = datasets.lgssm_config(1, 4, 5)
    keys = m.random.split(m.random.PRNGKey(2), 6).reshape(2, 3, 2)
    x0 = model.initial_sample(0.0, keys)
    assert x0.shape == (2, 3, 4)
    x_one = model.initial_sample(0.0, keys[1, 2])
    assert x_one.shape == (4,) and torch.equal(x_one, x0[1, 2])
    x1 = model.transition_sample(x0.reshape(6, 4), 0.0, 1.0, m.random.split(m.random.PRNGKey(3), 6))
    assert x1.shape == (6, 4)
    lik = model.likelihood_potential(x1, y[1], 1.0)
    p = opf.lgssm_params(model.initial_mean, model.initial_covariance, model.transition_matrix,
                         model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)
    want = opf.lik_potential(p, x1.cpu().numpy().astype(np.float64), y[1].cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(lik.cpu().numpy(), want, rtol=1e-5, atol=1e-5)
    # the bootstrap filter's per-particle rules are the model's
    bf = m.ssm.BootstrapFilter()
    assert torch.equal(bf.intermediate_log_weight(model, None, 0.0, x1, y[1], 1.0), -lik)
    sv = m.ssm.StochasticVolatility(1, 0.0, 1.0, 0.95, 0.04, 0.6)
    xs = torch.linspace(-2, 2, 9, device="cuda").reshape(9, 1)
    q = opf.sv_params(0.95, 0.2, np.sqrt(0.6))
    want = opf.lik_potential(q, xs.cpu().numpy().astype(np.float64), np.array([0.3]))
    np.testing.assert_allclose(sv.likelihood_potential(xs, torch.tensor([0.3]), 0.0).cpu().numpy(), want,
                               rtol=1e-5, atol=1e-5)
    # a step of length 0.5: the noise scales with sqrt(0.5) (stochastic_volatility.py:86-93)
    ks = m.random.split(m.random.PRNGKey(0), 9)
    x_half = sv.transition_sample(xs, 0.0, 0.5, ks).cpu().numpy().astype(np.float64)
    eps = np.stack([orng.normal(k, (1,)) for k in ks.cpu().numpy()]).astype(np.float64)
    want = opf.transition(q, xs.cpu().numpy().astype(np.float64), eps, np.sqrt(0.5))
    np.testing.assert_allclose(x_half, want, rtol=1e-5, atol=1e-6)


def test_models_without_a_kernel_raise_the_documented_error(m):
    class MyModel(m.ssm.StateSpaceModel):
        dim = dim_obs = 1

    class MyFilter(m.ssm.BootstrapFilter):
        def proposal_sample(self, *args):
            return None

    keys = m.random.split(m.random.PRNGKey(0), 4)
    with pytest.raises(NotImplementedError, match="compiled models only"):
        m.ssm.BootstrapFilter().initial_sample_and_weight_vectorised(MyModel(), None, 0.0, keys)
    with pytest.raises(NotImplementedError, match="compiled models only"):
        MyModel().initial_sample(0.0, keys)
    sv = m.ssm.StochasticVolatility(1, 0.0, 1.0, 0.9, 0.1, 0.5)
    with pytest.raises(NotImplementedError, match="compiled models only"):
        MyFilter().propose_and_intermediate_weight_vectorised(sv, torch.zeros((4, 1)), 0.0, torch.zeros(1), 1.0, keys)
