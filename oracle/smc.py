"""Adaptive tempered SMC (RM)MetropolisedSMCSampler restated in NumPy (test infrastructure only).

Follows mocat/src/transport/smc.py step by step (key tree included):
  startup   TransportSampler.startup (mocat/src/transport/sampler.py:18-34): (key, sub) = split(key),
            particles = prior_sample(split(sub, n)[i]); TemperedSMCSampler.startup (smc.py:128-164):
            key = split(key, 2n + 1)[-1]; MetropolisedSMCSampler.startup (smc.py:267-296): first
            temperature by regula falsi, lw = -b1 * lik, logZ = lse(lw) - log n, then the vmapped
            Underdamped.startup: key = split(key, 3)[0]
  iterate   SMCSampler.update (smc.py:73-99): keys = split(key, n + 2), key = keys[-1], resample with
            keys[-2] if ESS <= thr * n, forward_proposal (smc.py:337-365) per particle with keys[i]:
            Underdamped.startup (split(keys[i], 3)[0], re-evaluation at the current temperature), then
            mcmc_steps Metropolis updates (oracle.mcmc.iterate); adapt (smc.py:193-217, 406-421).
  regula_falsi  mocat/src/utils.py:205-237 (called "bisect" there), FP64.
The likelihood potential of the moved particles is the tracked one (not (U - prior) / beta).
Pinned by: mocat's own bisect tests (mocat/src/tests/test_utils.py:178-194) and the closed-form
Gaussian evidence (the method of mocat/src/tests/test_transport.py:49-55).
"""
import numpy as np

from . import mcmc, pf, rng


def regula_falsi(fun, b0, b1, max_iter=1000, tol=1e-5):
    e0, e1 = fun(b0), fun(b1)
    increasing = e1 > e0
    it = 0
    while not (min(abs(e0), abs(e1)) < tol or it >= max_iter or (e0 < 0 and e1 < 0) or (e0 > 0 and e1 > 0)
               or np.isnan(e0) or np.isnan(e1)):
        npos = b0 - e0 * (b1 - b0) / (e1 - e0)
        ne = fun(npos)
        if (ne > 0) if increasing else (ne < 0):
            b1, e1 = npos, ne
        else:
            b0, e0 = npos, ne
        it += 1
    return (b1 if abs(e1) < abs(e0) else b0), (b0, b1), (e0, e1), it


def next_temperature(lw, lik, beta, ess, retain, max_temperature=1.0, tol=1e-5, max_iter=1000):
    lw = np.asarray(lw, np.float64)
    lik = np.asarray(lik, np.float64)
    thr = np.log(ess * retain)
    f = lambda b: pf.log_ess(lw - (b - beta) * lik) - thr
    return regula_falsi(f, beta, max_temperature, max_iter, tol)


def run(pot_grad, prior_sample, d, n, key, stepsize, mcmc_steps=1, leapfrog_steps=1, retain=0.9,
        thr_resample=0.5, resampling="systematic", rm_stepsize=None, target=0.651, max_iter=1000, dtype=np.float64):
    """pot_grad(x, beta) -> (U, grad, prior, lik). Returns dict of per-iteration temperature, logZ, ess
    and the final particles / log weights."""
    key = np.asarray(key, np.uint32)
    key, sub = rng.split(key, 2)
    x = np.stack([prior_sample(k) for k in rng.split(sub, n)]).astype(dtype)
    key = rng.split(key, 2 * n + 1)[-1]
    _, _, pr, lik = pot_grad(x, 0.0)
    beta, _, _, _ = next_temperature(np.zeros(n), lik, 0.0, float(n), retain)
    lw = -beta * np.asarray(lik, np.float64)
    lnc = pf.lse(lw) - np.log(n)
    ess = np.exp(pf.log_ess(lw))
    key = rng.split(key, 3)[0]
    h = float(stepsize)
    temps, lncs, esss = [beta], [lnc], [ess]
    mom = np.zeros((n, d), dtype)
    it = 0
    while beta < 1.0 and it < max_iter:
        it += 1
        keys_all = rng.split(key, n + 2)
        key = keys_all[-1]
        if ess <= thr_resample * n:
            w = np.exp(lw - pf.lse(lw)).astype(np.float32)
            a = pf.resample_indices(w, keys_all[-2], resampling)
            x, mom = x[a], mom[a]
            lw = np.zeros(n)
            ess = float(n)
        pkeys = rng.split_batch(keys_all[:n], 3)[:, 0]
        U, g, pr, lik = pot_grad(x, beta)
        st = dict(value=x, momenta=mom, potential=U, grad_potential=g, prior_potential=pr,
                  likelihood_potential=lik, stepsize=np.full(n, h, dtype))
        alphas = []
        fn = lambda z: pot_grad(z, beta)
        for s in range(mcmc_steps):
            st, pkeys, al, acc = mcmc.iterate(st, pkeys, 1, fn, leapfrog_steps, np.inf, "mh", dtype=dtype)
            alphas.append(al)
        x, mom = st["value"], st["momenta"]
        lik = np.asarray(st["likelihood_potential"], np.float64)
        bnew, _, _, _ = next_temperature(lw, lik, beta, ess, retain)
        lw_new = lw - (bnew - beta) * lik
        lnc += pf.lse(lw_new) - pf.lse(lw)
        ess = np.exp(pf.log_ess(lw_new))
        if rm_stepsize is not None:
            wts = np.exp(lw_new - lw_new.max())
            am = np.average(np.mean(alphas, axis=0), weights=wts)
            h = float(np.exp(np.log(h) + rm_stepsize * (am - target)))
        lw = lw_new
        beta = bnew
        temps.append(beta)
        lncs.append(lnc)
        esss.append(ess)
    return dict(temperature=np.array(temps), log_norm_constant=np.array(lncs), ess=np.array(esss), value=x,
                log_weight=lw, stepsize=h)


def gaussian_pot_grad(mean, var, prior_var):
    mean = np.asarray(mean, np.float64)
    lam = 1.0 / np.asarray(var, np.float64)
    prec = 1.0 / prior_var
    d = mean.shape[0]

    def f(x, beta):
        x = np.asarray(x, np.float64)
        df = x - mean
        lik = 0.5 * np.sum(lam * df * df, axis=-1)
        pr = 0.5 * prec * np.sum(x * x, axis=-1) + 0.5 * d * (np.log(2 * np.pi) - np.log(prec))
        g = prec * x + beta * lam * df
        return pr + beta * lik, g, pr, lik
    return f


def gaussian_log_evidence(mean, var, prior_var, beta):
    """log int N(x; 0, prior_var I) exp(-beta * 0.5 sum (x - m)^2 / var) dx (closed form, diagonal)."""
    mean = np.asarray(mean, np.float64)
    lam = 1.0 / np.asarray(var, np.float64)
    b = np.asarray(beta, np.float64)[..., None]
    a = 1.0 + b * lam * prior_var
    return np.sum(-0.5 * np.log(a) - 0.5 * b * lam * mean ** 2 / a, axis=-1)
