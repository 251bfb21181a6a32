#!/usr/bin/env python3
"""Benchmark: batched HMC on Bayesian logistic regression, one MI355X.

Workload: X 2000 x 50 ~ N(0,1), theta* ~ N(0, I), y ~ Bernoulli(sigmoid(X theta*)), prior N(0, 10 I);
8192 independent chains of Underdamped(leapfrog_steps=10, friction=inf) = HMC with Robbins-Monro
stepsize adaptation to 0.651 acceptance; FP32 state, bf16x3 MFMA contractions.
A step is one HMC iteration of all chains; the K timed steps run as one launch of the fused kernel.

value = ESS/s of the whole 2000-iteration chain, whatever --steps is: per chain, mocat's
ess_autocorrelation of the potential trace over iterations 1001-2000, summed over chains, divided by
the wall time of all 2000 iterations (one launch, started from the prior).

Also reported: the error of the split-precision contraction against the FP64 engine at this shape,
the roofline of the fused kernel (algorithmic FLOPs 4*n*d*L per chain-iteration over its average
time per iteration from HIP events on the launch stream, against the 2.5 PFLOP/s dense bf16 MFMA
peak), the same chains as unfused FP32 torch ops on the same GPU, and a CPU baseline (the oracle's
NumPy port of the chain on the host cores: 64 chains of the whole 2000 iterations, ESS/s measured on
those CPU chains).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_ROWS, DIM, CHAINS, LEAPFROG, ITERATIONS = 2000, 50, 8192, 10, 2000
PEAK_TFLOPS = {"bf16x3": 2500.0, "fp32": 157.3}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hmc_lr_bf16x3_c8192_n2000_d50_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--chains", type=int, default=CHAINS)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--cpu-baseline", dest="cpu", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--cpu-chains", type=int, default=64)
    ap.add_argument("--no-plain", dest="plain", action="store_false", default=True,
                    help="skip the unfused FP32 torch chain on the same GPU")
    ap.add_argument("--iters", type=int, default=500, help="iterations of the svgd4 workload")
    ap.add_argument("--mode", default="hmc", choices=["hmc", "copy", "pf1", "pf5", "smc3", "svgd4", "ffbsi", "nlg"],
                    help="hmc: the headline workload. copy: a plain 256 MiB device copy (calibration run for HBM "
                         "byte counters). pf1 / pf5 / smc3 / svgd4 / ffbsi / nlg: the other workloads, one JSON line each")
    return ap.parse_args()


def copy_calibration():
    """Five device-to-device copies of 256 MiB (read 256 MiB + write 256 MiB each)."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy", "bytes_read_per_copy": src.numel() * 4, "bytes_written_per_copy": dst.numel() * 4}))


def clone(c):
    import torch
    from mocat_amd import cdict
    out = cdict()
    for k, v in c.__dict__.items():
        if isinstance(v, torch.Tensor):
            out.__dict__[k] = v.clone()
        elif isinstance(v, cdict):
            out.__dict__[k] = clone(v)
        else:
            out.__dict__[k] = v
    return out


def plain_fp32_run(X, y, state, extra, iters, prec=0.1):
    """The same HMC chain (same start, keys and RM adaptation) as unfused FP32 torch ops on the same
    GPU: GEMM, elementwise, GEMM per leapfrog step, one op at a time. Returns (wall seconds of the
    `iters` iterations, potential trace [iters][C])."""
    import torch
    import mocat_amd.random as mr
    Xt = torch.as_tensor(X, device="cuda")
    yt = torch.as_tensor(y, device="cuda")
    th = state.value.clone()
    g = state.grad_potential.clone()
    U = state.potential.clone()
    h = state.stepsize.clone()
    keys = extra.random_key.clone()
    C, d = th.shape
    trace = torch.empty((iters, C), dtype=torch.float32, device="cuda")

    def pot_grad(t):
        z = t @ Xt.T
        r = torch.sigmoid(z) - yt
        lik = (torch.nn.functional.softplus(z) - yt * z).sum(1)
        gr = prec * t + r @ Xt
        return 0.5 * prec * (t * t).sum(1) + lik, gr

    def one(it):
        nonlocal th, g, U, h, keys
        s = mr.split(keys, 2)
        p = mr.normal(s[:, 1].contiguous(), (d,))
        k = mr.split(s[:, 0].contiguous(), LEAPFROG + 1)[:, 0].contiguous()
        x, gg, pp = th, g, p
        K0 = 0.5 * (p * p).sum(1)
        for _ in range(LEAPFROG):
            ph = pp - (h[:, None] / 2) * gg
            x = x + h[:, None] * ph
            U1, gg = pot_grad(x)
            pp = ph - (h[:, None] / 2) * gg
        a = torch.clamp(torch.exp(-U1 + U - 0.5 * (pp * pp).sum(1) + K0), max=1.0).nan_to_num(0.0)
        s = mr.split(k, 2)
        keys = s[:, 0].contiguous()
        u = mr.uniform(s[:, 1].contiguous(), (1,))[:, 0]
        acc = u < a
        th = torch.where(acc[:, None], x, th)
        g = torch.where(acc[:, None], gg, g)
        U = torch.where(acc, U1, U)
        h = torch.exp(torch.log(h) + it ** (-2 / 3) * (a - 0.651))
        trace[it - 1] = U

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(1, iters + 1):
        one(it)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, trace


def cpu_baseline(X, y, chains, iters):
    """The oracle's NumPy port of the same HMC chains on the host cores (bounded sample: `chains`
    chains of the full `iters`-iteration chain). ESS/s is measured on these CPU chains exactly as the
    headline: ESS of the potential trace over the second half, over the wall time of all iterations."""
    import numpy as np
    import torch
    from oracle import mcmc as omcmc, rng as orng
    import mocat_amd as m
    pg = omcmc.lr_pot_grad_fn(X, y, 0.1)
    st, keys = omcmc.startup(orng.PRNGKey(0), chains, DIM, lambda k: np.sqrt(10.0) * orng.normal(k, (DIM,)), pg)
    trace = np.empty((iters, chains), np.float32)
    t0 = time.perf_counter()
    for it in range(1, iters + 1):
        st, keys, a, acc = omcmc.iterate(st, keys, it, pg, LEAPFROG, np.inf, "rm", 0.651)
        trace[it - 1] = st["potential"]
    dt = time.perf_counter() - t0
    ess = float(m.metrics.ess_autocorrelation(torch.tensor(trace[iters // 2:], device="cuda")).sum().item())
    return {"value": ess / dt, "unit": "ESS/s", "cores": _host_threads(), "kind": "port",
            "sample": f"{chains} chains x {iters} iterations of the NumPy oracle ({dt:.1f} s); ESS of the potential "
                      f"over iterations {iters // 2 + 1}-{iters} of these CPU chains / wall time of all iterations",
            "ess_total": ess, "chain_iterations_per_s": chains * iters / dt}


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def _host_threads():
    try:
        from threadpoolctl import threadpool_info
        return int(max([i.get("num_threads", 1) for i in threadpool_info()] or [1]))
    except Exception:
        return len(os.sched_getaffinity(0))


def _cpu_sample(fn, units, unit, sample):
    """Times fn() (the NumPy oracle on a bounded sample of the workload) on the host cores."""
    t0 = time.perf_counter()
    fn()
    dt = time.perf_counter() - t0
    return {"value": units / dt, "unit": unit, "cores": _host_threads(), "kind": "port",
            "sample": f"{sample} ({dt:.1f} s)"}


def bench_pf1():
    """Linear-Gaussian SSM d = 4, T = 100, N = 4096, systematic resampling every step; also 256
    independent filters in one launch. Unit: particle-steps/s."""
    import mocat_amd as m
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, 100)
    pf = m.ssm.BootstrapFilter()
    run1 = lambda: m.ssm.run_particle_filter_for_marginals(model, pf, y, t, m.random.PRNGKey(0), n=4096,
                                                           ess_threshold=1.0, resampling="systematic")
    run1()
    out, dt = _timed(run1)
    keys = m.random.split(m.random.PRNGKey(1), 256)
    yn = y.cpu().numpy()
    runb = lambda: m.ssm.run_particle_filters([model] * 256, [yn] * 256, [k for k in keys], [4096] * 256, 1.0,
                                              "systematic")
    runb()
    res, dtb = _timed(runb)
    # scaled variant for bandwidth evidence: one filter of 2^20 particles (chunked path, no trajectory)
    big = 1 << 20
    runs = lambda: m.ssm.run_particle_filters([model], [yn], [keys[0]], [big], 1.0, "systematic")
    runs()
    _, dts = _timed(runs)
    steps = big * 100
    from oracle import pf as opf, rng as orng
    p = opf.lgssm_params(model.initial_mean, model.initial_covariance, model.transition_matrix,
                         model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)
    cpu = _cpu_sample(lambda: opf.run_pf(p, yn, orng.PRNGKey(0), 4096, 1.0, "systematic"), 4096 * 100,
                      "particle-steps/s", "oracle bootstrap filter on the whole workload: 4096 particles x 100 steps, "
                      "systematic resampling every step")
    print(json.dumps({"mode": "pf1", "metric": "particle-steps/s", "cpu_baseline": cpu,
                      "single_filter": 4096 * 100 / dt,
                      "single_filter_ms": dt * 1e3, "batch256": 256 * 4096 * 100 / dtb, "batch256_ms": dtb * 1e3,
                      "log_likelihood": out.log_likelihood.item(), "store_trajectory": True,
                      "scaled_n2e20": {"ms": dts * 1e3, "particle_steps_per_s": steps / dts,
                                       "algorithmic_GBps": 48 * steps / dts / 1e9,
                                       "hbm_frac": 48 * steps / dts / 8e12, "bytes_per_particle_step": 48,
                                       "store_trajectory": False}}))


def bench_pf5():
    """256 stochastic-volatility filters, particles log-uniform 256..65536, lengths 50..1000, one launch."""
    import mocat_amd as m
    from mocat_amd import datasets
    models, ys, ns = datasets.sv_batch_config(0)
    keys = m.random.split(m.random.PRNGKey(0), len(models))
    # the schedule autotuner times the candidate schedules (chunked-path threshold x launch mode) on
    # this batch once, untimed, and records its choice; the timed runs use the recorded schedule
    from mocat_amd.ssm import schedule as sched
    t0 = time.perf_counter()
    run = lambda: m.ssm.run_particle_filters(models, ys, [k for k in keys], ns, 1.0, "systematic", schedule="auto")
    run()
    tune_s = time.perf_counter() - t0
    choice = sched.lookup(sched.signature(models, ys, ns, "systematic", False, False, None))
    res, dt = _timed(run)
    n_large = sum(1 for v in ns if v >= choice["large_n"])
    launches = (3 * (max(len(y) for y, v in zip(ys, ns) if v >= choice["large_n"]) - 1) + 3) if n_large else 0
    if n_large and choice["launch"] == "persistent":
        launches = 3  # init, the persistent kernel, the abort check
    E = sum(n * len(y) for n, y in zip(ns, ys))
    big = max(n * len(y) for n, y in zip(ns, ys))
    rung = lambda: m.ssm.run_particle_filters(models, ys, [k for k in keys], ns, 1.0, "systematic", grad=True,
                                              schedule="auto")
    rung()
    _, dtg = _timed(rung)
    import numpy as np
    from oracle import pf as opf, rng as orng
    work = [n * len(y) for n, y in zip(ns, ys)]
    order = list(np.argsort(work))
    pick, tot = [], 0
    for i in order[len(order) // 2:]:  # filters from the median size up, until ~6e6 particle-steps
        pick.append(int(i))
        tot += work[i]
        if tot >= 6e6:
            break

    def cpu_run():
        for i in pick:
            md = models[i]
            q = opf.sv_params(md.phi, md.sigma, md.beta, md.initial_mean, md.initial_covariance)
            opf.run_pf(q, np.asarray(ys[i]), orng.PRNGKey(i), ns[i], 1.0, "systematic")

    cpu = _cpu_sample(cpu_run, tot, "particle-steps/s",
                      f"oracle SV filters {pick} of the batch ({tot} particle-steps, from the median filter size up)")
    gchoice = sched.lookup(sched.signature(models, ys, ns, "systematic", True, False, None))
    print(json.dumps({"mode": "pf5", "metric": "particle-steps/s", "cpu_baseline": cpu, "value": E / dt,
                      "ms": dt * 1e3,
                      "schedule": {"chosen": {k: choice[k] for k in ("large_n", "launch")},
                                   "timings_ms": choice["timings_ms"], "tuning_seconds": tune_s,
                                   "chosen_with_gradient": {k: gchoice[k] for k in ("large_n", "launch")},
                                   "filters_on_chunked_path": n_large,
                                   "chunked_path_kernel_launches": launches,
                                   "chunked_path_submissions": 1 if choice["launch"] == "graph" else launches},
                      "with_gradient": E / dtg, "with_gradient_ms": dtg * 1e3,
                      "particle_steps": E, "largest_filter_particle_steps": big,
                      "bytes_per_particle_step": 24, "achieved_GBps_algorithmic": 24 * E / dt / 1e9}))


def bench_smc3(n_particles=65536):
    """Tempered SMC on Bayesian logistic regression, X 20000 x 100, 5 MALA moves per temperature."""
    import mocat_amd as m
    from mocat_amd import datasets
    X, y, _ = datasets.logistic_regression(0, 20000, 100)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision="bf16x3")
    smp = m.RMMetropolisedSMCSampler(m.Underdamped(stepsize=0.02), mcmc_steps=5, ess_threshold_retain=0.5,
                                     ess_threshold_resample=1.0, resampling="systematic")
    run = lambda: m.run(sc, smp, n_particles, m.random.PRNGKey(0), stack=())
    run()  # untimed: code-object loading and first allocations
    chain, dt = _timed(run)
    n_temp = int(chain.temperature.shape[0]) - 1
    evals = 6 * n_particles * n_temp
    import numpy as np
    from oracle import mcmc as omcmc, rng as orng
    tonp = lambda a: (a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)).astype(np.float64)
    Xn, yn = tonp(X), tonp(y)
    pg = omcmc.lr_pot_grad_fn(Xn, yn, 0.1)
    st, keys = omcmc.startup(orng.PRNGKey(0), 1024, 100, lambda k: np.sqrt(10.0) * orng.normal(k, (100,)), pg)

    def cpu_run():
        nonlocal st, keys
        for it in range(1, 11):
            st, keys, _, _ = omcmc.iterate(st, keys, it, pg, 1, np.inf, "rm", 0.651)

    cpu = _cpu_sample(cpu_run, 1024 * 10, "particle-moves/s",
                      "oracle RM-MALA moves of 1024 particles x 10 on the same X (20000 x 100); temperature search and "
                      "resampling not included")
    # stated error: the same run (same keys) with every contraction in FP64, and the Monte Carlo spread
    # of the bf16x3 estimator over two more seeds
    import torch
    lz = chain.log_norm_constant[-1].item()
    sc64 = m.LogisticRegression(X, y, prior_variance=10.0, precision="fp64")
    chain64, dt64 = _timed(lambda: m.run(sc64, smp, n_particles, m.random.PRNGKey(0), stack=()))
    lz64 = chain64.log_norm_constant[-1].item()
    seeds = [lz] + [m.run(sc, smp, n_particles, m.random.PRNGKey(s), stack=()).log_norm_constant[-1].item()
                    for s in (1, 2)]
    xs = chain.final_state.value if hasattr(chain, "final_state") else None
    err = None
    if xs is not None:
        _, lp, _, gl = sc.evaluate(xs)
        _, lp64, _, gl64 = sc64.evaluate(xs)
        err = {"lik_potential_rel_max": float(((lp - lp64).abs() / lp64.abs().clamp_min(1e-30)).max().item()),
               "grad_lik_rel_max": float(((gl - gl64).abs().amax(1) / gl64.abs().amax(1).clamp_min(1e-30))
                                         .max().item()),
               "at": "final particles of the bf16x3 run, FP64 engine as reference"}
    print(json.dumps({"mode": "smc3", "metric": "particle-moves/s", "cpu_baseline": cpu,
                      "value": 5 * n_particles * n_temp / dt,
                      "seconds": dt, "temperatures": n_temp, "log_evidence": lz,
                      "log_evidence_fp64_run": lz64, "temperatures_fp64_run": int(chain64.temperature.shape[0]) - 1,
                      "fp64_run_seconds": dt64,
                      "log_evidence_err_vs_fp64": abs(lz - lz64),
                      "log_evidence_rel_err_vs_fp64": abs(lz - lz64) / abs(lz64),
                      "log_evidence_seeds_0_1_2": seeds, "log_evidence_mc_sd": float(np.std(seeds, ddof=1)),
                      "contraction_error_vs_fp64": err,
                      "gradient_evals_per_s": evals / dt, "algorithmic_TFLOPs": evals * 4 * 20000 * 100 / dt / 1e12}))


def bench_svgd4(n=4096, d=100, iters=500):
    """SVGD, 4-component diagonal Gaussian mixture in d = 100, median bandwidth, adagrad 1e-2."""
    import numpy as np
    import mocat_amd as m
    rs = np.random.default_rng(0)
    means = rs.normal(0, 3, (4, d))
    var = rs.uniform(0.5, 2.0, (4, d))
    w = rs.dirichlet(np.ones(4))
    sc = m.GaussianMixture(means, var, w)
    # a short run of the same shapes first: code-object loading and first allocations stay untimed
    m.run(sc, m.SVGD(stepsize=1e-2, max_iter=10, bandwidth_update="median"), n, m.random.PRNGKey(0))
    smp = m.SVGD(stepsize=1e-2, max_iter=iters, bandwidth_update="median")
    chain, dt = _timed(lambda: m.run(sc, smp, n, m.random.PRNGKey(0)))
    flop = 6.0 * n * n * d * iters
    quality = svgd_quality(m, sc, smp, chain)
    from oracle import svgd as osvgd
    nc = 1024
    xc = rs.normal(0, 1, (nc, d))
    gsq, mom = np.zeros((nc, d)), np.zeros((nc, d))

    def cpu_run():
        nonlocal xc, gsq, mom
        for _ in range(8):
            _, g = osvgd.gmm_potential_grad(xc, means, var, w)
            h = osvgd.median_bandwidth(xc)
            xc, gsq, mom = osvgd.adagrad(xc, gsq, mom, -osvgd.phi_gemm(xc, g, h), 1e-2)

    cpu = _cpu_sample(cpu_run, 8, "iterations/s", f"oracle SVGD, {nc} particles x 8 iterations (median bandwidth, "
                      "mixture gradient, Stein direction, adagrad); the workload's 4096 particles cost ~16x per iteration")
    print(json.dumps({"mode": "svgd4", "metric": "iterations/s", "cpu_baseline": cpu, "value": iters / dt,
                      "seconds": dt,
                      "stein_algorithmic_TFLOPs_incl_bandwidth_and_potential": flop / dt / 1e12,
                      "final_bandwidth": chain.bandwidth[-1].item(), **quality}))


def svgd_quality(m, sc, smp, chain):
    """Untimed checks on the final particles: the Stein direction of one more update (bf16x3 MFMA
    path) against an FP64 evaluation of the same formula, and the KSD (Gaussian kernel at the final
    bandwidth) after the first and the last iteration, both as mocat's ksd computes it (grad U where
    a Stein kernel takes the score) and with the score -grad U (the textbook discrepancy)."""
    import torch
    st, ex = clone(chain.final_state), clone(chain.final_extra)
    x = st.value.to(torch.float64)
    g = st.grad_potential.to(torch.float64)
    h = float(chain.bandwidth[-1].item())
    phi = smp.step(st, ex, sc).to(torch.float64)
    n = x.shape[0]
    sq = (x * x).sum(1)
    D = (sq[:, None] + sq[None, :] - 2.0 * x @ x.T).clamp_min(0.0)
    K = torch.exp(-D / (2.0 * h * h))
    ref = (-(K @ g) + (K.sum(1, keepdim=True) * x - K @ x) / (h * h)) / n
    err = float(((phi - ref).norm() / ref.norm()).item())
    kern = m.kernels.Gaussian(bandwidth=h)
    x0 = m.run(sc, m.SVGD(stepsize=1e-2, max_iter=1, bandwidth_update="median"), n, m.random.PRNGKey(0))
    first = x0.final_state  # one update from the same start: its potential gradient is at hand
    last = chain.final_state
    return {"stein_direction_rel_err_vs_fp64": err,
            "mean_potential": {"after_first_iteration": float(first.potential.mean().item()),
                               "final": float(last.potential.mean().item())},
            "ksd_mocat_convention": {"after_first_iteration": m.ksd(first, kern), "final": m.ksd(last, kern)},
            "ksd_score": {"after_first_iteration": m.ksd(first.value, kern, grad_potential=-first.grad_potential),
                          "final": m.ksd(last.value, kern, grad_potential=-last.grad_potential)}}


def bench_ffbsi(n=4096, T=100):
    """Backward simulation after the pf1 filter (LGSSM d = 4, T = 100): n filter particles and n
    trajectories, full resampling (n x n transition potentials per step)."""
    import mocat_amd as m
    from mocat_amd import datasets
    model, y, t = datasets.lgssm_config(0, 4, T)
    marg = m.ssm.run_particle_filter_for_marginals(model, m.ssm.BootstrapFilter(), y, t, m.random.PRNGKey(0), n=n,
                                                   ess_threshold=1.0, resampling="systematic")
    run = lambda: m.ssm.backward_simulation(model, marg, m.random.PRNGKey(1), n)
    run()
    res, dt = _timed(run)
    runr = lambda: m.ssm.backward_simulation(model, marg, m.random.PRNGKey(1), n, maximum_rejections=32)
    runr()
    rr, dtr = _timed(runr)
    pairs = (T - 1) * n * n
    from oracle import ffbsi as offbsi, rng as orng
    nc, tc = 1024, 10
    vals = marg.value[:tc, :nc].double().cpu().numpy()
    lws = marg.log_weight[:tc, :nc].double().cpu().numpy()
    pot = offbsi.lgssm_transition(model.transition_matrix, model.transition_covariance)
    cpu = _cpu_sample(lambda: offbsi.backward_simulation_full(pot, vals, lws, orng.PRNGKey(1), nc),
                      (tc - 1) * nc * nc, "pair evaluations/s",
                      f"oracle full backward simulation, {nc} particles x {nc} trajectories x {tc} steps")
    print(json.dumps({"mode": "ffbsi", "metric": "pair evaluations/s", "cpu_baseline": cpu, "value": pairs / dt,
                      "ms": dt * 1e3,
                      "particles": n, "trajectories": n, "T": T, "ms_per_backward_step": dt * 1e3 / (T - 1),
                      "note": "per pair: one transition potential (d = 4) and half a threefry block",
                      "rejection_32": {"ms": dtr * 1e3, "ms_per_backward_step": dtr * 1e3 / (T - 1),
                                       "density_evals_per_step": float(rr.num_transition_evals[1:].mean().item())}}))


def bench_nlg(T=100):
    """Optimal-proposal particle filter and ensemble Kalman filter on the pf1 linear-Gaussian model
    (d = 4, T = 100): the optimal filter with 4096 particles (trajectory stored) and 2^20 (chunked path),
    the EnKF with 2^20 members (every ensemble stored). Unit: particle-steps/s. CPU baseline: the
    NumPy oracle's EnKF on 131072 members x 30 steps."""
    import numpy as np
    import mocat_amd as m
    from mocat_amd import datasets
    from oracle import nlg as onlg, rng as orng
    model, y, t = datasets.lgssm_config(0, 4, T)
    yn = y.cpu().numpy()
    opt = m.ssm.OptimalNonLinearGaussianParticleFilter()
    run1 = lambda: m.ssm.run_particle_filter_for_marginals(model, opt, y, t, m.random.PRNGKey(0), n=4096,
                                                           resampling="systematic")
    run1()
    o1, dt1 = _timed(run1)
    big = 1 << 20
    runb = lambda: m.ssm.run_particle_filters([model], [yn], [m.random.PRNGKey(1)], [big], 0.5, "systematic",
                                              particle_filter=opt)
    runb()
    _, dtb = _timed(runb)
    enkf = m.ssm.EnsembleKalmanFilter()
    rune = lambda: m.ssm.run_particle_filter_for_marginals(model, enkf, y, t, m.random.PRNGKey(2), n=big)
    rune()
    oe, dte = _timed(rune)
    ps = big * T
    # CPU: the oracle's EnKF (FP64 NumPy) on a bounded sample
    p = onlg.startup(model.initial_mean, model.initial_covariance, model.transition_matrix,
                     model.transition_covariance, model.likelihood_matrix, model.likelihood_covariance)
    nc, tc = 131072, 30
    t0 = time.perf_counter()
    onlg.run_enkf(p, yn[:tc], orng.PRNGKey(2), nc)
    dtc = time.perf_counter() - t0
    print(json.dumps({"mode": "nlg", "metric": "particle-steps/s",
                      "optimal_pf_4096": {"particle_steps_per_s": 4096 * T / dt1, "ms": dt1 * 1e3,
                                          "log_likelihood_y1T_given_y0": o1.log_likelihood.item(),
                                          "store_trajectory": True},
                      "optimal_pf_n2e20": {"particle_steps_per_s": ps / dtb, "ms": dtb * 1e3,
                                           "ms_per_step": dtb * 1e3 / T, "store_trajectory": False},
                      "enkf_n2e20": {"particle_steps_per_s": ps / dte, "ms": dte * 1e3,
                                     "ms_per_step": dte * 1e3 / (T - 1), "store_trajectory": True,
                                     "bytes_per_particle_step": 48,
                                     "algorithmic_GBps": 48 * big * (T - 1) / dte / 1e9,
                                     "hbm_frac": 48 * big * (T - 1) / dte / 8e12},
                      "cpu_baseline": {"value": nc * tc / dtc, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                                       "sample": f"oracle EnKF, {nc} members x {tc} steps ({dtc:.1f} s)"}}))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only (independent chains and filters shard trivially across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    if a.mode == "copy":
        copy_calibration()
        return
    if a.mode != "hmc":
        if a.mode == "svgd4":
            bench_svgd4(iters=a.iters)
        else:
            {"pf1": bench_pf1, "pf5": bench_pf5, "smc3": bench_smc3, "ffbsi": bench_ffbsi,
             "nlg": bench_nlg}[a.mode]()
        return
    import mocat_amd as m
    from mocat_amd import datasets

    X, y, _ = datasets.logistic_regression(0, N_ROWS, DIM)
    sc = m.LogisticRegression(X, y, prior_variance=10.0, precision=a.precision)
    sampler = m.Underdamped(leapfrog_steps=LEAPFROG)
    extra = m.cdict(random_key=m.random.PRNGKey(0))
    state, extra = sampler.startup(sc, ITERATIONS, None, extra, n_chains=a.chains, correction=m.RMMetropolis())
    evs = lambda: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))

    # untimed warm-up iterations (first launches) on a copy of the start state
    if a.warmup > 0:
        ws, we = clone(state), clone(extra)
        sampler.iterate(sc, ws, we, a.warmup)
        torch.cuda.synchronize()
        del ws, we

    # the timed steps: K iterations of all chains from the start state, one launch of the fused kernel
    K = a.steps
    ks, ke = clone(state), clone(extra)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sampler.iterate(sc, ks, ke, K)
    torch.cuda.synchronize()
    wall_steps = time.perf_counter() - t0
    del ks, ke

    # the headline: config 2's whole 2000-iteration chain from the start state, one launch, potential
    # trace on the device; ESS over iterations 1001-2000 / wall time of all 2000 iterations
    hs, he = clone(state), clone(extra)
    trace_pot = torch.empty((ITERATIONS, a.chains), dtype=torch.float32, device="cuda")
    trace_alpha = torch.empty((ITERATIONS, a.chains), dtype=torch.float32, device="cuda")
    ev0, ev1 = evs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    sampler.iterate(sc, hs, he, ITERATIONS, {"potential": trace_pot, "alpha": trace_alpha})
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / ITERATIONS  # HIP events on the launch stream

    half = trace_pot[ITERATIONS // 2:]
    ess_total = float(m.metrics.ess_autocorrelation(half).sum().item())
    value = ess_total / wall
    acc = float(trace_alpha[ITERATIONS // 2:].mean().item())

    # stated error of the split-precision contraction at this shape: the final states of the chain
    # evaluated by the bf16x3 engine and by the FP64 engine
    ref = m.LogisticRegression(X, y, prior_variance=10.0, precision="fp64")
    _, lp, _, gl = sc.evaluate(hs.value)
    _, lp64, _, gl64 = ref.evaluate(hs.value)
    gscale = gl64.abs().amax(1).clamp_min(1e-30)
    err = {"reference": "FP64 engine (same final states of the 2000-iteration chain)",
           "lik_potential_rel_max": float(((lp - lp64).abs() / lp64.abs().clamp_min(1e-30)).max().item()),
           "grad_lik_rel_max": float(((gl - gl64).abs().amax(1) / gscale).max().item())}

    flop_per_launch = 4.0 * N_ROWS * DIM * LEAPFROG * a.chains  # per iteration of all chains
    achieved = flop_per_launch / (kernel_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[a.precision]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_iteration")

    out = {
        "metric": "ESS/s (batched HMC, Bayesian logistic regression)",
        "value": value,
        "unit": "ESS/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": wall_steps / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+f32" if a.precision == "bf16x3" else "f32",
        "data": "synthetic (X ~ N(0,1), seed 0)",
        "config": {"workload": "batched HMC on Bayesian logistic regression", "n": N_ROWS, "d": DIM,
                   "chains": a.chains, "leapfrog_steps": LEAPFROG, "friction": "inf",
                   "correction": "RMMetropolis(target=0.651)", "precision": a.precision,
                   "iterations": ITERATIONS, "ess_window": f"iterations {ITERATIONS // 2 + 1}-{ITERATIONS}"},
        "ess_total": ess_total,
        "acceptance": acc,
        "chain_wall_s": wall,
        "chain_iterations_per_s": a.chains * ITERATIONS / wall,
        "error_vs_fp64": err,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "kernel": "hmc_lr_bf16_kernel<64,512>" if a.precision == "bf16x3"
                     else "hmc_lr_kernel<Fp32Engine<64,512>>",
                     "kernel_ms_per_iteration": kernel_ms, "flop_per_iteration": flop_per_launch,
                     "iterations_per_launch": ITERATIONS,
                     "issued_mfma_flop_per_iteration": (3 * 4.0 * 64 * (math.ceil(N_ROWS / 32) * 32) * LEAPFROG
                                                        * a.chains) if a.precision == "bf16x3" else None},
    }
    if a.plain:
        pw, ptrace = plain_fp32_run(X, y, state, extra, ITERATIONS)
        pess = float(m.metrics.ess_autocorrelation(ptrace[ITERATIONS // 2:]).sum().item())
        out["plain_fp32_gpu"] = {"value": pess / pw, "unit": "ESS/s", "ms_per_iteration": pw / ITERATIONS * 1e3,
                                 "speedup_ess_per_s": value / (pess / pw),
                                 "speedup_per_iteration": (pw / ITERATIONS) / (wall / ITERATIONS),
                                 "note": "the same chains as unfused FP32 torch GEMM + elementwise ops, same GPU, "
                                         f"{ITERATIONS} iterations"}
    if a.cpu:
        out["cpu_baseline"] = cpu_baseline(X, y, a.cpu_chains, ITERATIONS)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
