"""mocat's Underdamped (HMC / MALA) chain with Metropolis / Robbins-Monro corrections, restated in
NumPy over a batch of independent chains (test infrastructure only).

Follows, step by step and key by key:
  MCMCSampler.update            mocat/src/mcmc/sampler.py:92-111   (iter += 1; always; proposal; correct)
  Underdamped.always            mocat/src/mcmc/standard_mcmc.py:107-123
  Underdamped.proposal          mocat/src/mcmc/standard_mcmc.py:125-143
  utils._leapfrog               mocat/src/utils.py:108-146
  Underdamped.acceptance_prob.  mocat/src/mcmc/standard_mcmc.py:145-153
  Metropolis.correct            mocat/src/mcmc/metropolis.py:48-70
  RMMetropolis.correct/adapt    mocat/src/mcmc/metropolis.py:118-161 (log update, monotonicity -1)
Arithmetic is FP32 like the reference (jax x64 is never enabled); `dtype=np.float64` gives the
FP64 run used for error statements.
Pinned by: the leapfrog known-answer test of mocat/src/tests/test_utils.py:121-147 (bit-exact
in f32), the RNG pinning of oracle.rng, and posterior moments within Monte Carlo error.
"""
import numpy as np

from . import rng

LOG2PI = np.log(2 * np.pi)


def leapfrog(grad_fn, x, p, g, stepsize, steps, dtype=np.float32):
    """Returns (x, p, g, pot) after `steps` velocity-Verlet steps; grad_fn(x) -> (pot, grad)."""
    h = dtype(stepsize)
    x = np.asarray(x, dtype)
    p = np.asarray(p, dtype)
    g = np.asarray(g, dtype)
    pot = None
    for _ in range(steps):
        ph = (p - (h / dtype(2.0)) * g).astype(dtype)
        x = (x + h * ph).astype(dtype)
        pot, g = grad_fn(x)
        pot = np.asarray(pot, dtype)
        g = np.asarray(g, dtype)
        p = (ph - (h / dtype(2.0)) * g).astype(dtype)
    return x, p, g, pot


def kinetic(p, dtype):
    d = p.shape[-1]
    return (dtype(0.5) * np.sum(p * p, axis=-1) + dtype(d / 2 * LOG2PI)).astype(dtype)


def chain_keys(random_key, n_chains):
    return rng.split(random_key, n_chains)


def iterate(state, keys, it, pot_grad, leapfrog_steps, friction, correction="rm", target=0.651,
            rm_scale=1.0, rm_neg_exponent=2.0 / 3.0, adapt_cut_off=np.inf, dtype=np.float32):
    """One MCMCSampler.update for each chain.

    state: dict of arrays with leading chain axis: value, momenta, potential, grad_potential,
    prior_potential, likelihood_potential, stepsize (per chain). keys: (C, 2) uint32.
    it: value of extra.iter after the increment (1-based). pot_grad(x) -> (U, grad, prior, lik).
    Returns (new_state, new_keys, alpha, accepted).
    """
    C, d = state["value"].shape
    s = rng.split_batch(keys, 2)
    k, sub = s[:, 0], s[:, 1]
    xi = rng.normal_batch(sub, d).astype(dtype)
    k = rng.split_batch(k, leapfrog_steps + 1)[:, 0]
    us = np.zeros(C, np.float32)
    if correction in ("mh", "rm"):
        s = rng.split_batch(k, 2)
        k, sub2 = s[:, 0], s[:, 1]
        us = rng.uniform_batch(sub2, 1)[:, 0]
    new_keys = k
    h = state["stepsize"].astype(dtype)
    fr = dtype(friction)
    with np.errstate(invalid="ignore", over="ignore"):
        decay = np.exp(-fr * h).astype(dtype)
        noise = np.sqrt(dtype(1) - np.exp(dtype(-2) * fr * h)).astype(dtype)
    p0 = ((-state["momenta"]).astype(dtype) * decay[:, None] + noise[:, None] * xi).astype(dtype)

    x = state["value"].astype(dtype)
    g = state["grad_potential"].astype(dtype)
    p = p0
    hh = h[:, None]
    res = None
    for _ in range(leapfrog_steps):
        ph = (p - (hh / dtype(2.0)) * g).astype(dtype)
        x = (x + hh * ph).astype(dtype)
        res = pot_grad(x)
        g = np.asarray(res[1], dtype)
        p = (ph - (hh / dtype(2.0)) * g).astype(dtype)
    U1, _, pr1, lk1 = [np.asarray(r, dtype) for r in res]
    p_prop = -p
    U0 = state["potential"].astype(dtype)
    with np.errstate(over="ignore", invalid="ignore"):
        a = np.minimum(dtype(1), np.exp(-U1 + U0 - kinetic(p_prop, dtype) + kinetic(p0, dtype))).astype(dtype)
    a = np.where(np.isnan(a), dtype(0), a).astype(dtype)
    if correction in ("mh", "rm"):
        acc = us < a
    else:
        acc = np.ones(C, bool)
        a = np.ones(C, dtype)
    out = {}
    out["value"] = np.where(acc[:, None], x, state["value"]).astype(dtype)
    out["grad_potential"] = np.where(acc[:, None], g, state["grad_potential"]).astype(dtype)
    out["momenta"] = np.where(acc[:, None], p_prop, p0).astype(dtype)
    out["potential"] = np.where(acc, U1, U0).astype(dtype)
    out["prior_potential"] = np.where(acc, pr1, state["prior_potential"]).astype(dtype)
    out["likelihood_potential"] = np.where(acc, lk1, state["likelihood_potential"]).astype(dtype)
    stepsize = state["stepsize"].astype(dtype)
    if correction == "rm" and it <= adapt_cut_off:
        rm = dtype(rm_scale) * np.power(dtype(it), dtype(-rm_neg_exponent), dtype=dtype)
        stepsize = np.exp(np.log(stepsize) + rm * (a - dtype(target))).astype(dtype)
    out["stepsize"] = stepsize
    return out, new_keys, a, acc


def lr_pot_grad_fn(X, y, prec, temperature=1.0, dtype=np.float32):
    from . import lr
    def f(x):
        U, g, pp, lp = lr.potential_and_grad(x, X, y, prec, temperature)
        return U.astype(dtype), g.astype(dtype), pp.astype(dtype), lp.astype(dtype)
    return f


def startup(random_key, n_chains, d, prior_sample, pot_grad, stepsize=None, dtype=np.float32):
    """Batched MCMCSampler/Underdamped/RMMetropolis startup: chain c uses split(key, C)[c] and
    mocat's single-chain startup (mocat/src/mcmc/sampler.py:32-61, standard_mcmc.py:86-105,
    metropolis.py:91-116). prior_sample(key) -> (d,) array."""
    keys = rng.split(random_key, n_chains)
    vals = np.empty((n_chains, d), dtype)
    for c in range(n_chains):
        k, sub = rng.split(keys[c], 2)
        vals[c] = prior_sample(sub)
        k = rng.split(k, 3)[0]
        keys[c] = k
    U, g, pp, lp = pot_grad(vals)
    h = dtype(2.38 ** 2 / d) if stepsize is None else dtype(stepsize)
    state = dict(value=vals, momenta=np.zeros((n_chains, d), dtype), potential=np.asarray(U, dtype),
                 grad_potential=np.asarray(g, dtype), prior_potential=np.asarray(pp, dtype),
                 likelihood_potential=np.asarray(lp, dtype), stepsize=np.full(n_chains, h, dtype))
    return state, keys
