#!/usr/bin/env python
"""Benchmark of the Riemannian integrator hot path on one MI355X.

Headline workload (the metric's "4096 chains, tuned vs plain"): Bayesian
logistic regression with the Fisher metric G = X^T diag(s(1-s)) X + I/100,
X [2048, 64] iid N(0, 1/64), beta* ~ N(0, I), y ~ Bernoulli(sigmoid(X beta*)),
data seed 2; RMHMC with the implicit midpoint, step 0.1, 10 steps, fixed point
capped at 6 iterations; 4096 chains started at N(0, I) (seed 1).

A "step" is one transition of all 4096 chains (10 integrator steps each).

* tuned form (the reported ``value``): float32 state, Gram and quadratic-form
  contractions on bf16 MFMA (split operands, float32 accumulation), float32
  sweep inverse with one refinement step, float64 log-det / energies /
  accept test / fixed-point norm; persistent launches of several transitions
  with device Philox noise.
* plain form: the float64 kernel, one launch per transition from the host
  (the reference's per-step API).  It starts from the tuned form's warmed-up
  chain state, so both time stationary sampling.

``value`` = chains * 10 * steps / wall seconds of the timed region (inputs
resident in HBM).  min-ESS/s uses the multi-chain ESS of
geomjax/diagnostics.py:78-209 over a fixed window of post-warm-up transitions
(100 for the 2048 x 64 posterior, 40 for the uneven batch; at least 20
warm-up transitions), independent of --steps, divided by the window's own wall
time; the tuned and the plain form run that window from the same warmed state
with the same device draws, so their min-ESS estimates are paired.  Extra fields report the
same comparison at 1024 chains, for 256 logistic posteriors of uneven size in
one launch (8192 and 4096 chains), for Neal's funnel D=100 (RMHMC, NUTS-RMHMC
and LMC with the Monge metric, 8192 chains), and the cost of the tuning
derivatives (tangent launch vs primal transition).  The CPU baseline is the float64 oracle
restatement (dense metric + autodiff, as the reference) timed on this host's
cores on a bounded sample of the headline workload.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import functools
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "min-ESS/sec & leapfrog-steps/sec, 4096 chains, tuned vs plain GPU form"
# logistic regression (headline)
N_ROWS, D_LOG, CHAINS, EPS_LOG, L_LOG, MAXIT_LOG, PRIOR, DATA_SEED, INIT_SEED = 2048, 64, 4096, 0.1, 10, 6, 100.0, 2, 1
CHUNK_LOG = 10  # transitions per persistent launch (tuned form)
WARM_MIN = 20  # untimed warm-up transitions at least (whatever --warmup says)
# min-ESS/s windows: fixed numbers of post-warm-up transitions, independent of --steps
ESS_WINDOW_LOG, ESS_WINDOW_BATCH = 100, 40
# Neal's funnel (secondary)
D_F, CHAINS_F, EPS_F, L_F, MAXIT_F, ALPHA2, SIGMA, SEED_F = 100, 8192, 0.03, 32, 8, 1.0, 3.0, 1
CHUNK_F = 100
BF16_PEAK = 2.5e15
FP32_PEAK = 157.3e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "logistic_rmhmc_f32_kernel_logistic_2048x64_4096chains.md")


def measured_traffic():
    """HBM bytes per launch of the headline kernel, as measured by the FETCH_SIZE and
    WRITE_SIZE counters (calibrated on a copy kernel) in the committed profile summary;
    None if the summary is absent.  The counters need a profiler run of their own, so
    this figure is not re-measured by a plain bench run."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def mfma_flops_logistic(num_fevals_total):
    """Algorithmic MFMA work of RMHMC-Fisher on logistic regression: per evaluation
    of the vector field the Gram X^T Lambda X (2 N D^2) and X G^-1 for the trace
    term (2 N D^2)."""
    return num_fevals_total * 4 * N_ROWS * D_LOG * D_LOG


def flops_rmhmc_monge_funnel(num_fevals_total, chains, transitions):
    """Algorithmic flops of RMHMC-Monge on the funnel (transcendentals count 1):
    per evaluation of the vector field 33 D + 25 flops + 1 exp; per chain-step
    11 D; per transition two energies (8 D + 2 log) and the rank-one draw (10 D)."""
    return (num_fevals_total * (33 * D_F + 26) + chains * transitions * L_F * 11 * D_F
            + chains * transitions * (18 * D_F + 2))


def min_ess(x):
    """min over coordinates of the multi-chain ESS of [T, C, D] samples (chunked over D)."""
    import torch

    from geomjax_amd.diagnostics import effective_sample_size

    out = []
    for d0 in range(0, x.shape[2], 8):
        out.append(effective_sample_size(x[:, :, d0:d0 + 8], chain_axis=1, sample_axis=0))
    ess = torch.cat([o.reshape(-1) for o in out])
    return float(ess.min().item()), int(ess.argmin().item())


def cpu_baseline(seconds_budget=30.0):
    """Oracle (float64 dense Fisher metric + autodiff restatement) on this host's cores."""
    import torch

    from geomjax_amd.synthetic import make_logistic
    from oracle import models as om
    from oracle import rmhmc as orm

    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    M = om.LogisticRegression(X, y, PRIOR)
    kern = orm.build_kernel(max_iters=MAXIT_LOG)
    g = torch.Generator().manual_seed(INIT_SEED)
    C = 64
    s = orm.init(torch.randn(C, D_LOG, generator=g, dtype=torch.float64), M.logp)
    # untimed warm-up on a few chains
    kern((torch.randn(2, D_LOG, generator=g, dtype=torch.float64), torch.rand(2, generator=g, dtype=torch.float64)),
         orm.init(s.position[:2], M.logp), M.logp, EPS_LOG, M.fisher_metric_fn, L_LOG)
    n_tr, t_used = 0, 0.0
    while True:
        z = torch.randn(C, D_LOG, generator=g, dtype=torch.float64)
        u = torch.rand(C, generator=g, dtype=torch.float64)
        t0 = time.perf_counter()
        s, _ = kern((z, u), s, M.logp, EPS_LOG, M.fisher_metric_fn, L_LOG)
        t_used += time.perf_counter() - t0
        n_tr += 1
        if t_used > seconds_budget / 2 or n_tr >= 3:
            break
    return {"value": C * L_LOG * n_tr / t_used, "unit": "leapfrog-steps/s", "cores": threads, "kind": "port",
            "sample": f"{C} chains x {n_tr} transitions of the headline logistic workload (float64 dense "
                      f"Fisher metric, Cholesky + autodiff restatement, {t_used:.1f} s)"}


def _warm(state, model, key, counter, W, integ, **kw):
    """Untimed tuned-form warm-up of at least WARM_MIN transitions (the first
    call also autotunes the launch shape)."""
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import sample

    for _ in range(max(WARM_MIN, W) // CHUNK_LOG):
        r = sample(gr.PRNGKey(key.seed, counter), state, model, EPS_LOG, _fisher(), L_LOG, CHUNK_LOG, integrator=integ,
                   store_samples=False, **kw)
        state, counter = r.state, counter + CHUNK_LOG
    return state, counter


def _fisher():
    import geomjax_amd as gj

    return gj.FisherMetric()


def paired_ess_window(model, warmed, key, integ, window, plain, ess_fn):
    """min-ESS/s of the tuned and the plain form over the same fixed window of
    `window` transitions after warm-up, both started from the same warmed
    state and fed the same device draws (Philox streams keyed by chain and
    transition), so the two min-ESS estimates are paired and their ratio is not
    dominated by the estimator's noise.  min-ESS/s = min-ESS of the window's
    samples / the window's wall time."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import build_kernel, sample

    C, D = warmed.position.shape
    dev = warmed.position.device
    base = 500_000
    samples = torch.empty(window, C, D, dtype=torch.float32, device=dev)
    state = warmed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(window // CHUNK_LOG):
        r = sample(gr.PRNGKey(key.seed, base + c * CHUNK_LOG), state, model, EPS_LOG, _fisher(), L_LOG, CHUNK_LOG,
                   integrator=integ, samples_out=samples[c * CHUNK_LOG:(c + 1) * CHUNK_LOG])
        state = r.state
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ess = ess_fn(samples)
    del samples
    out = {"window_transitions": window, "draws": "paired (same Philox streams)",
           "tuned": {"min_ess": ess, "min_ess_per_sec": ess / wall, "wall_s": wall, "ms_per_step": wall / window * 1e3}}
    if plain:
        kern = build_kernel(integ)
        s64 = gj.rmhmc.init(warmed.position.double(), model)
        s64, _ = kern(gr.PRNGKey(key.seed, base - 1), s64, model, EPS_LOG, _fisher(), L_LOG)  # untimed: first launch
        s64 = gj.rmhmc.init(warmed.position.double(), model)
        psamples = torch.empty(window, C, D, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(window):
            s64, _ = kern(gr.PRNGKey(key.seed, base + t), s64, model, EPS_LOG, _fisher(), L_LOG)
            psamples[t].copy_(s64.position)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pess = ess_fn(psamples)
        # how closely the paired chains track each other at the window's end
        drift = float((psamples[-1] - state.position.double()).abs().max().item())
        del psamples
        out["plain"] = {"min_ess": pess, "min_ess_per_sec": pess / pwall, "wall_s": pwall,
                        "ms_per_step": pwall / window * 1e3, "value": C * L_LOG * window / pwall,
                        "unit": "leapfrog-steps/s", "form": "float64 kernel, one launch per transition"}
        out["max_abs_state_difference_at_window_end"] = drift
        out["tuned_vs_plain_min_ess_per_sec"] = out["tuned"]["min_ess_per_sec"] / out["plain"]["min_ess_per_sec"]
    return out


# waves per chain the library uses for each width class when the field leaves it at 0
# (kDefaultWavesPerChain of the C library: two for D <= 32, one for 33..64)
DEFAULT_WAVES = {32: 2, 64: 1}


def waves_for_width(field, D):
    """Waves per chain of the tuned kernel for a posterior of D <= 64 features
    under the C ABI's waves_per_chain field (4 bits per width class)."""
    cls = 32 if D <= 32 else 64
    w = (field >> (0 if cls == 32 else 4)) & 0xF
    return w or DEFAULT_WAVES[cls]


def logistic_forms(chains, K, W, dev, plain=True):
    """Tuned and plain RMHMC-Fisher on the 2048 x 64 logistic posterior with `chains` chains."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import autotune
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import implicit_midpoint, sample
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    model, metric = gj.LogisticRegression(X, y, PRIOR, device=dev), gj.FisherMetric()
    integ = functools.partial(implicit_midpoint, max_iters=MAXIT_LOG)
    q_init = torch.randn(chains, D_LOG, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    key = gr.key(INIT_SEED)
    K = max(CHUNK_LOG, (K // CHUNK_LOG) * CHUNK_LOG)

    # ---- tuned: warm-up, then the timed region (leapfrog-steps/s) ----
    state, counter = _warm(gj.rmhmc.init(q_init.to(dev, torch.float32), model), model, key, 10_000, W, integ)
    warmed = state
    fevs, accs = [], []
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for c in range(K // CHUNK_LOG):
        r = sample(gr.PRNGKey(key.seed, counter), state, model, EPS_LOG, metric, L_LOG, CHUNK_LOG,
                   integrator=integ, store_samples=False)
        state, counter = r.state, counter + CHUNK_LOG
        fevs.append(r.num_fevals)
        accs.append(r.acceptance_rate)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = ev0.elapsed_time(ev1) / 1e3
    fev_total = int(torch.stack(fevs).sum().item())
    n_launch = K // CHUNK_LOG
    waves = autotune.logistic_waves_field(model, chains)
    tuned = {"value": chains * L_LOG * K / wall, "unit": "leapfrog-steps/s", "ms_per_step": wall / K * 1e3,
             "acceptance_rate": float(torch.stack(accs).mean().item()),
             "fevals_per_chain_transition": fev_total / chains / K, "dtype": "f32 state, bf16x3 MFMA, f64 energies",
             "launches": n_launch, "avg_launch_ms": kernel_s / n_launch * 1e3,
             "mfma_flops_per_launch": mfma_flops_logistic(fev_total) / n_launch, "steps": K,
             "waves_per_chain": waves_for_width(waves, D_LOG)}
    log(f"logistic tuned, {chains} chains: {tuned['value']:.3e} leapfrog-steps/s, {tuned['ms_per_step']:.2f} ms/step, "
        f"acc {tuned['acceptance_rate']:.3f}, fevals/chain/tr {tuned['fevals_per_chain_transition']:.1f}, "
        f"waves per chain {tuned['waves_per_chain']} (autotuned)")

    win = paired_ess_window(model, warmed, key, integ, ESS_WINDOW_LOG, plain, lambda x: min_ess(x)[0])
    tuned["min_ess_per_sec"] = win["tuned"]["min_ess_per_sec"]
    tuned["min_ess"] = win["tuned"]["min_ess"]
    res = {"tuned": tuned, "plain": win.get("plain"), "ess_window": win}
    if plain:
        res["tuned_vs_plain"] = {"leapfrog_steps_per_sec": tuned["value"] / win["plain"]["value"],
                                 "min_ess_per_sec": win["tuned_vs_plain_min_ess_per_sec"]}
        log(f"logistic plain, {chains} chains: {win['plain']['value']:.3e} leapfrog-steps/s; "
            f"min-ESS/s tuned {win['tuned']['min_ess_per_sec']:.3e} plain {win['plain']['min_ess_per_sec']:.3e} "
            f"over {ESS_WINDOW_LOG} paired transitions; tuned/plain {res['tuned_vs_plain']}")
    return res


def batch_min_ess(samples, counts, dims):
    """min over posteriors and their coordinates of the multi-chain ESS (chains of one posterior)."""
    import torch

    from geomjax_amd.diagnostics import effective_sample_size

    out, c0 = [], 0
    for n, d in zip(counts, dims):
        ess = effective_sample_size(samples[:, c0:c0 + n, :d], chain_axis=1, sample_axis=0)
        out.append(ess.min())
        c0 += n
    return float(torch.stack(out).min().item())


def uneven_batch_forms(chains_per_posterior, K, W, dev, plain=True):
    """Tuned and plain RMHMC-Fisher on 256 logistic posteriors of uneven size in one launch.

    Posteriors as the uneven-batch configuration generates them: D drawn from
    {8, 16, 32, 64, 128}, N log-uniform in [256, 4096], data seed 3.  Chains of
    posteriors up to 64 features run one workgroup of one or two waves each
    (chosen per width by the autotuner), wider ones four waves."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import autotune
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import implicit_midpoint, sample
    from geomjax_amd.synthetic import make_uneven_batch

    data = make_uneven_batch(P=256, seed=3, dims=(8, 16, 32, 64, 128))
    models = [gj.LogisticRegression(X, y, PRIOR, device=dev) for X, y, _ in data]
    batch = gj.LogisticBatch(models, chains_per_posterior)
    counts, dims = batch.counts, [m.D for m in models]
    C, Dm = batch.num_chains, batch.dim
    integ = functools.partial(implicit_midpoint, max_iters=MAXIT_LOG)
    metric = gj.FisherMetric()
    q_init = torch.zeros(C, Dm, dtype=torch.float64)
    g = torch.Generator().manual_seed(INIT_SEED)
    c0 = 0
    for n, d in zip(counts, dims):
        q_init[c0:c0 + n, :d] = torch.randn(n, d, generator=g, dtype=torch.float64)
        c0 += n
    key = gr.key(INIT_SEED)
    K = max(CHUNK_LOG, (K // CHUNK_LOG) * CHUNK_LOG)
    work = sum(n * m.N * m.D * m.D for n, m in zip(counts, models))
    log(f"uneven batch: {len(models)} posteriors, {C} chains, sum chains*N*D^2 = {work:.3e}")

    state, counter = _warm(gj.rmhmc.init(q_init.to(dev, torch.float32), batch), batch, key, 10_000, W, integ)
    warmed = state
    fevs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(K // CHUNK_LOG):
        r = sample(gr.PRNGKey(key.seed, counter), state, batch, EPS_LOG, metric, L_LOG, CHUNK_LOG,
                   integrator=integ, store_samples=False)
        state, counter = r.state, counter + CHUNK_LOG
        fevs.append(r.num_fevals)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    waves = autotune.logistic_waves_field(batch, C)
    tuned = {"value": C * L_LOG * K / wall, "unit": "leapfrog-steps/s", "ms_per_step": wall / K * 1e3, "steps": K,
             "fevals_per_chain_transition": float(torch.stack(fevs).double().sum().item()) / C / K,
             "waves_per_chain": {"D<=32": waves_for_width(waves, 32), "33<=D<=64": waves_for_width(waves, 64),
                                 "65<=D<=128": 4, "chosen_by": "autotune.tune_logistic"}}
    log(f"uneven batch tuned: {tuned['value']:.3e} leapfrog-steps/s, {tuned['ms_per_step']:.2f} ms/step, "
        f"waves per chain {tuned['waves_per_chain']}")
    win = paired_ess_window(batch, warmed, key, integ, ESS_WINDOW_BATCH, plain,
                            lambda x: batch_min_ess(x, counts, dims))
    tuned["min_ess_per_sec"] = win["tuned"]["min_ess_per_sec"]
    tuned["min_ess"] = win["tuned"]["min_ess"]
    res = {"posteriors": len(models), "chains": C, "dims": "D in {8, 16, 32, 64, 128}, N in [256, 4096]",
           "tuned": tuned, "plain": win.get("plain"), "ess_window": win}
    if plain:
        res["tuned_vs_plain"] = {"leapfrog_steps_per_sec": tuned["value"] / win["plain"]["value"],
                                 "min_ess_per_sec": win["tuned_vs_plain_min_ess_per_sec"]}
        log(f"uneven batch plain: {win['plain']['value']:.3e} leapfrog-steps/s; tuned/plain {res['tuned_vs_plain']}")
    return res


def funnel_forms(K, W, dev, plain=True):
    """Tuned and plain RMHMC-Monge on Neal's funnel D=100 with 8192 chains.

    The tuned form runs as the configuration states it: each chunk of
    CHUNK_F transitions is one persistent launch captured with its noise draw
    and state copy in a HIP graph (GraphSampler), replayed per chunk; the
    launch shape (lanes per chain) comes from the autotuner."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import autotune
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import build_kernel, implicit_midpoint, sample
    from geomjax_amd.sampling import GraphSampler

    K = max(CHUNK_F, (K // CHUNK_F) * CHUNK_F)
    model, metric = gj.NealFunnel(D_F, sigma=SIGMA), gj.MongeMetric(ALPHA2)
    integ = functools.partial(implicit_midpoint, max_iters=MAXIT_F)
    q_init = torch.randn(CHAINS_F, D_F, generator=torch.Generator().manual_seed(SEED_F), dtype=torch.float64)
    key = gr.key(SEED_F)
    s32 = gj.rmhmc.init(q_init.to(dev, torch.float32), model)

    def run(lpc, state=s32, k=key, T=CHUNK_F, out=None):
        return sample(k, state, model, EPS_F, metric, L_F, T, integrator=integ, lanes_per_chain=lpc,
                      store_samples=out is not None, samples_out=out)

    lpc, timings = autotune.tune("rmhmc_monge_funnel", D_F, CHAINS_F, torch.float32, lambda l: run(l), repeats=2)
    log(f"funnel autotune lanes_per_chain: {lpc} ({ {k: round(v * 1e3, 2) for k, v in timings.items()} } ms)")
    counter, state = 10_000, s32
    for _ in range(max(1, W // CHUNK_F)):
        r = run(0, state, gr.PRNGKey(key.seed, counter))  # 0: the autotuned shape, as the product path takes it
        state, counter = r.state, counter + CHUNK_F
    # eager persistent launches (no graph), for comparison
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = run(0, state, gr.PRNGKey(key.seed, counter))
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3
    alg = gj.rmhmc(model, EPS_F, metric, L_F, integrator=integ)
    gs = GraphSampler(alg, state, CHUNK_F)  # captured once: noise draw + persistent launch + state copy
    gs.replay()
    samples = torch.empty(K, CHAINS_F, D_F, dtype=torch.float32, device=dev)
    fevs, accs = [], []
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for c in range(K // CHUNK_F):
        gs.replay()
        samples[c * CHUNK_F:(c + 1) * CHUNK_F].copy_(gs.samples)
        fevs.append(gs.num_fevals.clone())
        accs.append(gs.acceptance_rate.mean())
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = ev0.elapsed_time(ev1) / 1e3
    state = gs.state
    fev_total = int(torch.stack(fevs).sum().item())
    ess, _ = min_ess(samples)
    del samples
    n_launch = K // CHUNK_F
    flops = flops_rmhmc_monge_funnel(fev_total, CHAINS_F, K)
    tuned = {"value": CHAINS_F * L_F * K / wall, "unit": "leapfrog-steps/s", "min_ess_per_sec": ess / wall,
             "min_ess": ess, "acceptance_rate": float(torch.stack(accs).mean().item()), "lanes_per_chain": lpc,
             "kernel": f"funnel_rmhmc_kernel<float,{lpc},*,false>", "form": "HIP graph replays (GraphSampler)",
             "avg_replay_ms": kernel_s / n_launch * 1e3, "eager_persistent_launch_ms": eager_ms,
             "achieved_tflops": flops / kernel_s / 1e12, "frac_fp32_peak": flops / kernel_s / FP32_PEAK,
             "steps": K}
    log(f"funnel tuned: {tuned['value']:.3e} leapfrog-steps/s, min-ESS/s {tuned['min_ess_per_sec']:.3e}")
    res = {"tuned": tuned, "plain": None}
    if plain:
        s64 = gj.rmhmc.init(state.position.double(), model)
        kern = build_kernel(integ)
        s64, _ = kern(gr.PRNGKey(key.seed + 1, 0), s64, model, EPS_F, metric, L_F)
        psamples = torch.empty(K, CHAINS_F, D_F, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(K):
            s64, _ = kern(gr.PRNGKey(key.seed + 1, 1 + t), s64, model, EPS_F, metric, L_F)
            psamples[t].copy_(s64.position)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pess, _ = min_ess(psamples)
        del psamples
        res["plain"] = {"value": CHAINS_F * L_F * K / pwall, "unit": "leapfrog-steps/s",
                        "min_ess_per_sec": pess / pwall, "min_ess": pess, "dtype": "f64",
                        "form": "float64 kernel, one launch per transition"}
        res["tuned_vs_plain"] = {"leapfrog_steps_per_sec": tuned["value"] / res["plain"]["value"],
                                 "min_ess_per_sec": tuned["min_ess_per_sec"] / res["plain"]["min_ess_per_sec"]}
        log(f"funnel plain: {res['plain']['value']:.3e} leapfrog-steps/s; tuned/plain {res['tuned_vs_plain']}")
    return res


def funnel_lmc_forms(K, W, dev, plain=True):
    """Tuned and plain LMC-Monge (lmcmonge) on Neal's funnel D=100 with 8192 chains:
    inverse mass = ones (M = I), alpha2 = 1, step 0.03, 32 integrator steps."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import autotune
    from geomjax_amd import random as gr
    from geomjax_amd.lmcmonge import build_kernel as lmc_build_kernel
    from geomjax_amd.lmcmonge import init as lmc_init
    from geomjax_amd.lmcmonge import sample as lmc_sample

    K = max(CHUNK_F, (K // CHUNK_F) * CHUNK_F)
    model = gj.NealFunnel(D_F, sigma=SIGMA)
    q_init = torch.randn(CHAINS_F, D_F, generator=torch.Generator().manual_seed(SEED_F), dtype=torch.float64)
    key = gr.key(SEED_F)
    s32 = lmc_init(q_init.to(dev, torch.float32), model)
    im32 = torch.ones(D_F, dtype=torch.float32, device=dev)

    def run(lpc, state=s32, k=key, T=CHUNK_F, out=None):
        return lmc_sample(k, state, model, EPS_F, im32, L_F, ALPHA2, T, lanes_per_chain=lpc,
                               store_samples=out is not None, samples_out=out)

    lpc, timings = autotune.tune("lmcmonge_funnel", D_F, CHAINS_F, torch.float32, lambda l: run(l), repeats=2)
    log(f"funnel lmc autotune lanes_per_chain: {lpc} ({ {k: round(v * 1e3, 2) for k, v in timings.items()} } ms)")
    counter, state = 10_000, s32
    for _ in range(max(1, W // CHUNK_F)):
        r = run(lpc, state, gr.PRNGKey(key.seed, counter))
        state, counter = r.state, counter + CHUNK_F
    samples = torch.empty(K, CHAINS_F, D_F, dtype=torch.float32, device=dev)
    accs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(K // CHUNK_F):
        r = run(lpc, state, gr.PRNGKey(key.seed, counter), out=samples[c * CHUNK_F:(c + 1) * CHUNK_F])
        state, counter = r.state, counter + CHUNK_F
        accs.append(r.acceptance_rate)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ess, _ = min_ess(samples)
    del samples
    tuned = {"value": CHAINS_F * L_F * K / wall, "unit": "leapfrog-steps/s", "min_ess_per_sec": ess / wall,
             "min_ess": ess, "acceptance_rate": float(torch.stack(accs).mean().item()), "lanes_per_chain": lpc,
             "steps": K, "ms_per_launch": wall / (K // CHUNK_F) * 1e3}
    log(f"funnel lmc tuned: {tuned['value']:.3e} leapfrog-steps/s, min-ESS/s {tuned['min_ess_per_sec']:.3e}")
    res = {"tuned": tuned, "plain": None}
    if plain:
        s64 = lmc_init(state.position.double(), model)
        im64 = torch.ones(D_F, dtype=torch.float64, device=dev)
        kern = lmc_build_kernel()
        s64, _ = kern(gr.PRNGKey(key.seed + 1, 0), s64, model, EPS_F, im64, L_F, ALPHA2)
        psamples = torch.empty(K, CHAINS_F, D_F, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(K):
            s64, _ = kern(gr.PRNGKey(key.seed + 1, 1 + t), s64, model, EPS_F, im64, L_F, ALPHA2)
            psamples[t].copy_(s64.position)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pess, _ = min_ess(psamples)
        del psamples
        res["plain"] = {"value": CHAINS_F * L_F * K / pwall, "unit": "leapfrog-steps/s",
                        "min_ess_per_sec": pess / pwall, "min_ess": pess, "dtype": "f64",
                        "form": "float64 kernel, one launch per transition"}
        res["tuned_vs_plain"] = {"leapfrog_steps_per_sec": tuned["value"] / res["plain"]["value"],
                                 "min_ess_per_sec": tuned["min_ess_per_sec"] / res["plain"]["min_ess_per_sec"]}
        log(f"funnel lmc plain: {res['plain']['value']:.3e} leapfrog-steps/s; tuned/plain {res['tuned_vs_plain']}")
    return res


def nuts_forms(K, W, dev, static_plain=None, chunk=10, tuned=False):
    """NUTS with the RMHMC integrator (float64, or with ``tuned`` the float32
    state with float64 energies) on Neal's funnel D=100 with 8192 chains: Monge
    metric, the funnel workload's step size, max_num_doublings 10, "euc"
    criterion, device PRNG.  Timed as persistent launches of `chunk`
    transitions (each chain keeps its state on chip, so a launch lasts as long
    as the largest per-chain sum of trajectory lengths) and, for comparison,
    as one launch per transition (a launch lasts as long as its longest
    trajectory).  Reports integration steps/s, transitions/s and min-ESS/s,
    and min-ESS/s relative to the plain static-length RMHMC form when given."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.nuts import build_kernel as nuts_kernel
    from geomjax_amd.nuts import sample as nuts_sample
    from geomjax_amd.rmhmc import implicit_midpoint

    M = 10
    K = max(chunk, (K // chunk) * chunk)
    model, metric = gj.NealFunnel(D_F, sigma=SIGMA), gj.MongeMetric(ALPHA2)
    integ = functools.partial(implicit_midpoint, max_iters=MAXIT_F)
    kern = nuts_kernel(integ)
    dt = torch.float32 if tuned else torch.float64
    q_init = torch.randn(CHAINS_F, D_F, generator=torch.Generator().manual_seed(SEED_F), dtype=torch.float64)
    s = gj.rmhmc.init(q_init.to(dev).to(dt), model)
    r = nuts_sample(gr.PRNGKey(SEED_F + 7, 0), s, model, EPS_F, metric, max(W, chunk), max_num_doublings=M,
                    integrator=integ, store_samples=False)
    s, counter = r.state, max(W, chunk)
    warmed = s
    samples = torch.empty(K, CHAINS_F, D_F, dtype=dt, device=dev)
    steps, accs = [], []
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for c in range(K // chunk):
        r = nuts_sample(gr.PRNGKey(SEED_F + 7, counter), s, model, EPS_F, metric, chunk, max_num_doublings=M,
                        integrator=integ, samples_out=samples[c * chunk:(c + 1) * chunk])
        s, counter = r.state, counter + chunk
        steps.append(r.num_integration_steps)
        accs.append(r.acceptance_rate)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = ev0.elapsed_time(ev1) / 1e3
    st = torch.cat(steps)
    total_steps = int(st.long().sum().item())
    ess, _ = min_ess(samples)
    del samples
    res = {"value": total_steps / wall, "unit": "integration-steps/s", "transitions_per_sec": K / wall,
           "min_ess_per_sec": ess / wall, "min_ess": ess, "mean_integration_steps": total_steps / (K * CHAINS_F),
           "max_integration_steps": int(st.max().item()),
           "acceptance_rate": float(torch.cat(accs).mean().item()),
           "ms_per_transition": wall / K * 1e3, "kernel_ms_per_transition": kernel_s / K * 1e3,
           "kernel": f"funnel_nuts_kernel<2, {'float' if tuned else 'double'}>",
           "form": f"persistent launches of {chunk} transitions",
           "dtype": "f32 state, f64 energies" if tuned else "f64", "max_num_doublings": M, "steps": K}
    # the same chains, one launch per transition
    n1 = min(K, 10)
    s = warmed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(n1):
        s, info = kern(gr.PRNGKey(SEED_F + 7, 10 ** 6 + t), s, model, EPS_F, metric, M)
    torch.cuda.synchronize()
    res["one_launch_per_transition_ms"] = (time.perf_counter() - t0) / n1 * 1e3
    if static_plain is not None:
        res["min_ess_per_sec_vs_static_rmhmc_plain"] = res["min_ess_per_sec"] / static_plain["min_ess_per_sec"]
    log(f"funnel NUTS ({res['dtype']}): {res['value']:.3e} integration-steps/s, {res['mean_integration_steps']:.1f} steps per "
        f"transition, {res['ms_per_transition']:.2f} ms per transition persistent vs "
        f"{res['one_launch_per_transition_ms']:.2f} one launch each, min-ESS/s {res['min_ess_per_sec']:.3e}")
    return res


def logistic_nuts_forms(K, W, dev, chains=1024, integrator="rmhmc", tuned=False):
    """NUTS on the headline posterior (logistic regression 2048 x 64, Fisher
    metric, float64, one launch per transition) with the RMHMC integrator (the
    headline step size and fixed-point cap) or the general-metric LMC integrator
    (``integrator="lmc"``, same step size): max_num_doublings 5, "euc", device
    PRNG; with ``tuned`` the float32 state on the tuned kernel.  Reports
    integration steps/s and min-ESS/s."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.synthetic import make_logistic

    M = 5
    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    model = gj.LogisticRegression(X, y, PRIOR, device=dev)
    dt = torch.float32 if tuned else torch.float64
    q_init = 0.1 * torch.randn(chains, D_LOG, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    q_init = q_init.to(dt)
    if integrator == "lmc":
        from geomjax_amd.lmcmc import init as lmc_init
        from geomjax_amd.lmcmc_nuts import build_kernel as lnuts_kernel

        kname = "logistic_nuts_lmc_f32_kernel<2>" if tuned else "logistic_nuts_lmc_f64_kernel"
        kern, s = lnuts_kernel(), lmc_init(q_init.to(dev), model)
    else:
        from geomjax_amd.nuts import build_kernel as nuts_kernel
        from geomjax_amd.rmhmc import implicit_midpoint

        kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=MAXIT_LOG))
        kname = "logistic_nuts_f32_kernel<2>" if tuned else "logistic_nuts_f64_kernel"
        s = gj.rmhmc.init(q_init.to(dev), model)
    for t in range(W):
        s, _ = kern(gr.PRNGKey(INIT_SEED + 11, t), s, model, EPS_LOG, gj.FisherMetric(), M)
    samples = torch.empty(K, chains, D_LOG, dtype=dt, device=dev)
    steps, accs = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(K):
        s, info = kern(gr.PRNGKey(INIT_SEED + 11, W + t), s, model, EPS_LOG, gj.FisherMetric(), M)
        samples[t].copy_(s.position)
        steps.append(info.num_integration_steps)
        accs.append(info.acceptance_rate)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_steps = int(torch.stack(steps).long().sum().item())
    ess, _ = min_ess(samples)
    del samples
    res = {"value": total_steps / wall, "unit": "integration-steps/s", "min_ess_per_sec": ess / wall, "min_ess": ess,
           "mean_integration_steps": total_steps / (K * chains), "acceptance_rate": float(torch.stack(accs).mean()),
           "ms_per_transition": wall / K * 1e3, "chains": chains,
           "dtype": "f32 state, f64 energies" if tuned else "f64", "max_num_doublings": M, "steps": K,
           "kernel": kname}
    log(f"logistic NUTS ({integrator}, {res['dtype']}): {res['value']:.3e} integration-steps/s, {res['mean_integration_steps']:.1f} steps per "
        f"transition, min-ESS/s {res['min_ess_per_sec']:.3e}")
    return res


def logistic_lmc_forms(K, W, dev, chains=1024, L=10, tuned=False):
    """The general-metric LMC kernel (lmcmc, Fisher metric, one launch per
    transition) on the headline posterior (logistic 2048 x 64): step size 0.1, 10
    integration steps, device PRNG; float64, or with ``tuned`` the float32 state
    with split-bf16 MFMA Grams.  Reports leapfrog-steps/s, min-ESS/s and the
    algorithmic MFMA rate (per integration step five Gram matrices and one
    quadratic-form pass, 12 N D^2 flop; per transition one more of each, 4 N D^2)."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.lmcmc import build_kernel as lmc_kernel
    from geomjax_amd.lmcmc import init as lmc_init
    from geomjax_amd.synthetic import make_logistic

    dt = torch.float32 if tuned else torch.float64
    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    model = gj.LogisticRegression(X, y, PRIOR, device=dev)
    kern = lmc_kernel()
    q = 0.1 * torch.randn(chains, D_LOG, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    s = lmc_init(q.to(dev).to(dt), model)
    for t in range(W):
        s, _ = kern(gr.PRNGKey(INIT_SEED + 13, t), s, model, EPS_LOG, gj.FisherMetric(), L)
    samples = torch.empty(K, chains, D_LOG, dtype=dt, device=dev)
    accs = []
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for t in range(K):
        s, info = kern(gr.PRNGKey(INIT_SEED + 13, W + t), s, model, EPS_LOG, gj.FisherMetric(), L)
        samples[t].copy_(s.position)
        accs.append(info.acceptance_rate)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    ess, _ = min_ess(samples)
    del samples
    flops = (12.0 * L + 4.0) * N_ROWS * D_LOG ** 2 * chains * K
    res = {"value": chains * L * K / wall, "unit": "leapfrog-steps/s", "min_ess_per_sec": ess / wall, "min_ess": ess,
           "acceptance_rate": float(torch.stack(accs).mean()), "ms_per_transition": wall / K * 1e3, "chains": chains,
           "gpu_ms_per_transition": gpu_s / K * 1e3, "algorithmic_tflops": flops / gpu_s / 1e12,
           "num_integration_steps": L, "dtype": "f32 state, f64 energies" if tuned else "f64", "steps": K,
           "kernel": "logistic_lmc_f32_kernel<2, true>" if tuned else "logistic_lmc_f64_kernel"}
    log(f"logistic LMC (Fisher, {res['dtype']}): {res['value']:.3e} leapfrog-steps/s, "
        f"{res['ms_per_transition']:.2f} ms per transition, {res['algorithmic_tflops']:.1f} TFLOP/s algorithmic, "
        f"min-ESS/s {res['min_ess_per_sec']:.3e}")
    return res


def mmala_forms(K, W, dev, chains=1024, step_size=0.3):
    """Manifold MALA (Fisher metric, float64, corrected Metropolis test) on the
    headline posterior (logistic 2048 x 64): K transitions per chain in one
    persistent launch after W warm-up transitions, device PRNG.  Reports
    transitions/s, min-ESS/s and the acceptance rate."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    model = gj.LogisticRegression(X, y, PRIOR, device=dev)
    alg = gj.mmala(model, step_size, gj.FisherMetric(), reference_faithful=False)
    q = 0.1 * torch.randn(chains, D_LOG, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    s = alg.init(q.to(dev))
    if W:
        s = alg.sample_fn(gr.PRNGKey(INIT_SEED + 17, 0), s, W, store_samples=False).state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = alg.sample_fn(gr.PRNGKey(INIT_SEED + 17, W), s, K)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ess, _ = min_ess(res.samples)
    out = {"value": chains * K / wall, "unit": "transitions/s", "min_ess_per_sec": ess / wall, "min_ess": ess,
           "acceptance_rate": float(res.acceptance_rate.double().mean()), "ms_per_transition": wall / K * 1e3,
           "chains": chains, "step_size": step_size, "dtype": "f64", "steps": K,
           "kernel": "logistic_mmala_f64_kernel<false>"}
    log(f"MMALA (Fisher): {out['value']:.3e} transitions/s, acceptance {out['acceptance_rate']:.3f}, "
        f"min-ESS/s {out['min_ess_per_sec']:.3e}")
    return out


def dense_lmc_forms(dev, chains=1024, D=20, L=4, T=5, step_size=0.05):
    """LMC with a general dense user metric: Neal's funnel (D = 20) with the Monge
    metric I + g g^T written as a plain torch function and differentiated by
    torch.func; the integrator's linear algebra in the dense-metric kernels.
    Reports leapfrog-steps/s over T transitions of `chains` chains."""
    import math

    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr

    def logp(q):
        x, v = q[:-1], q[-1]
        return (-0.5 * v * v / 9.0 - 0.5 * math.log(2 * math.pi * 9.0)
                + torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi)))

    grad = torch.func.grad(logp)

    def metric(q):
        g = grad(q)
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + torch.outer(g, g)

    alg = gj.lmc(logp, step_size, metric, L)
    q = torch.randn(chains, D, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    s = alg.init(q.to(dev))
    s = alg.sample_fn(gr.PRNGKey(INIT_SEED + 19, 0), s, 1, store_samples=False).state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = alg.sample_fn(gr.PRNGKey(INIT_SEED + 19, 1), s, T, store_samples=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same chunk captured once into a HIP graph (no host dispatch per kernel)
    from geomjax_amd.sampling import GraphSampler

    gs = GraphSampler(alg, res.state, T, store_samples=False)
    gs.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 4
    for _ in range(reps):
        gs.replay()
    torch.cuda.synchronize()
    gwall = (time.perf_counter() - t0) / reps
    out = {"value": chains * L * T / gwall, "unit": "leapfrog-steps/s", "ms_per_transition": gwall / T * 1e3,
           "eager_ms_per_transition": wall / T * 1e3, "form": "HIP graph of 5 transitions, replayed",
           "acceptance_rate": float(res.acceptance_rate.double().mean()), "chains": chains, "dim": D,
           "num_integration_steps": L, "dtype": "f64", "steps": T,
           "kernels": "dense_lmc_half_step_kernel, dense_metric_position_kernel (+ torch.func for the user functions)"}
    log(f"dense-metric LMC (funnel D={D}, Monge by autodiff): {out['value']:.3e} leapfrog-steps/s, "
        f"{out['ms_per_transition']:.1f} ms per transition in a graph ({out['eager_ms_per_transition']:.1f} eager)")
    return out


def dense_rmhmc_forms(dev, chains=1024, D=20, L=4, T=3, step_size=0.05, max_iters=8):
    """RMHMC with a general dense user metric: Neal's funnel (D = 20) with the
    Monge metric I + g g^T as a plain torch function, implicit midpoint with at
    most 8 fixed-point iterations; T transitions captured in a HIP graph
    (the fixed point then always sweeps max_iters times) and, for comparison,
    run eagerly (early exit when every chain has converged)."""
    import math

    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import implicit_midpoint
    from geomjax_amd.sampling import GraphSampler

    def logp(q):
        x, v = q[:-1], q[-1]
        return (-0.5 * v * v / 9.0 - 0.5 * math.log(2 * math.pi * 9.0)
                + torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi)))

    grad = torch.func.grad(logp)

    def metric(q):
        g = grad(q)
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + torch.outer(g, g)

    alg = gj.rmhmc(logp, step_size, metric, L, integrator=functools.partial(implicit_midpoint, max_iters=max_iters))
    q = torch.randn(chains, D, generator=torch.Generator().manual_seed(INIT_SEED), dtype=torch.float64)
    s = alg.sample_fn(gr.PRNGKey(INIT_SEED + 23, 0), alg.init(q.to(dev)), 1, store_samples=False).state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = alg.sample_fn(gr.PRNGKey(INIT_SEED + 23, 1), s, T, store_samples=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gs = GraphSampler(alg, res.state, T, store_samples=False)
    gs.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        gs.replay()
    torch.cuda.synchronize()
    gwall = (time.perf_counter() - t0) / 3
    out = {"value": chains * L * T / gwall, "unit": "leapfrog-steps/s", "ms_per_transition": gwall / T * 1e3,
           "eager_ms_per_transition": wall / T * 1e3, "form": f"HIP graph of {T} transitions, replayed",
           "acceptance_rate": float(res.acceptance_rate.double().mean()),
           "fevals_per_chain_transition_eager": float(res.num_fevals.double().mean()) / T, "chains": chains,
           "dim": D, "num_integration_steps": L, "max_iters": max_iters, "dtype": "f64", "steps": T,
           "kernels": "dense_rmhmc_field_kernel (+ torch.func for the user functions)"}
    log(f"dense-metric RMHMC (funnel D={D}, Monge by autodiff): {out['value']:.3e} leapfrog-steps/s, "
        f"{out['ms_per_transition']:.1f} ms per transition in a graph ({out['eager_ms_per_transition']:.1f} eager)")
    return out


def lmc_alpha2_forms(dev, repeats=5):
    """Cost of the LMC-Monge alpha2 tangent (float64, the derivative ChEES-lmcmonge
    takes every iteration) against one float64 LMC-Monge transition on the funnel
    D=100 with 8192 chains, 32 integration steps."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.lmcmonge import build_kernel as lmc_kernel
    from geomjax_amd.lmcmonge import init as lmc_init
    from geomjax_amd.lmcmonge import proposal_alpha2_derivative

    model = gj.NealFunnel(D_F, sigma=SIGMA)
    q = torch.randn(CHAINS_F, D_F, generator=torch.Generator().manual_seed(SEED_F), dtype=torch.float64).to(dev)
    s = lmc_init(q, model)
    im = torch.ones(D_F, dtype=torch.float64, device=dev)
    kern = lmc_kernel()
    proposal_alpha2_derivative(gr.PRNGKey(3, 0), s, model, EPS_F, im, L_F, ALPHA2)
    kern(gr.PRNGKey(3, 0), s, model, EPS_F, im, L_F, ALPHA2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(repeats):
        proposal_alpha2_derivative(gr.PRNGKey(3, i), s, model, EPS_F, im, L_F, ALPHA2)
    torch.cuda.synchronize()
    t_tan = (time.perf_counter() - t0) / repeats
    t0 = time.perf_counter()
    for i in range(repeats):
        kern(gr.PRNGKey(3, i), s, model, EPS_F, im, L_F, ALPHA2)
    torch.cuda.synchronize()
    t_pri = (time.perf_counter() - t0) / repeats
    res = {"tangent_ms": t_tan * 1e3, "primal_f64_transition_ms": t_pri * 1e3, "tangent_over_primal": t_tan / t_pri,
           "chains": CHAINS_F, "dim": D_F, "num_integration_steps": L_F}
    log(f"lmc alpha2 tangent: {t_tan * 1e3:.2f} ms vs primal {t_pri * 1e3:.2f} ms ({t_tan / t_pri:.2f}x)")
    return res


def tangent_forms(dev, repeats=3):
    """Cost of the derivative of acceptance x squared jump w.r.t. (log eps, log alpha2):
    one tangent launch (one transition carried with its forward derivative) against one
    primal float64 transition, on the tuning workloads: funnel D=100 x 4096 chains and
    logistic 2048 x 64 x 1024 chains, 32 integrator steps, as many fixed-point
    iterations as the solver runs; the float32 fast path of each; and the
    float32 tangent of a 2048 x 128 posterior (256 chains), the width no float64
    tangent serves."""
    import torch

    import geomjax_amd as gj
    from geomjax_amd import random as gr
    from geomjax_amd.adjoint import esjd_acceptance
    from geomjax_amd.rmhmc import build_kernel, implicit_midpoint
    from geomjax_amd.synthetic import make_logistic

    out = {}
    cases = [("funnel_d100_4096_chains", gj.NealFunnel(D_F, sigma=SIGMA), gj.MongeMetric(ALPHA2), 4096, D_F, 0.03,
              MAXIT_F)]
    X, y, _ = make_logistic(N_ROWS, D_LOG, DATA_SEED)
    cases.append(("logistic_2048x64_1024_chains", gj.LogisticRegression(X, y, PRIOR, device=dev), gj.FisherMetric(),
                  1024, D_LOG, EPS_LOG, MAXIT_LOG))
    for name, model, metric, C, D, eps, maxit in cases:
        integ = functools.partial(implicit_midpoint, max_iters=maxit)
        q = 0.5 * torch.randn(C, D, generator=torch.Generator().manual_seed(7), dtype=torch.float64)
        s = gj.rmhmc.init(q.to(dev), model)
        kern = build_kernel(integ)
        esjd_acceptance(gr.PRNGKey(1, 0), s, model, eps, metric, 32, integrator=integ)
        kern(gr.PRNGKey(1, 0), s, model, eps, metric, 32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(repeats):
            r = esjd_acceptance(gr.PRNGKey(2, i), s, model, eps, metric, 32, integrator=integ)
        torch.cuda.synchronize()
        t_tan = (time.perf_counter() - t0) / repeats
        t0 = time.perf_counter()
        for i in range(repeats):
            kern(gr.PRNGKey(2, i), s, model, eps, metric, 32)
        torch.cuda.synchronize()
        t_pri = (time.perf_counter() - t0) / repeats
        out[name] = {"tangent_ms": t_tan * 1e3, "primal_f64_transition_ms": t_pri * 1e3,
                     "tangent_over_primal": t_tan / t_pri, "num_integration_steps": 32,
                     "mean_objective": float(r.objective.mean()), "mean_d_objective": r.d_objective.mean(0).tolist()}
        log(f"tangent {name}: {t_tan * 1e3:.2f} ms vs primal {t_pri * 1e3:.2f} ms ({t_tan / t_pri:.2f}x)")
        # the float32 fast path on the same positions
        s32 = gj.rmhmc.init(q.to(dev, torch.float32), model)
        esjd_acceptance(gr.PRNGKey(1, 0), s32, model, eps, metric, 32, integrator=integ)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(repeats):
            r32 = esjd_acceptance(gr.PRNGKey(2, i), s32, model, eps, metric, 32, integrator=integ)
        torch.cuda.synchronize()
        t32 = (time.perf_counter() - t0) / repeats
        out[name]["tangent_f32_fast_path_ms"] = t32 * 1e3
        out[name]["mean_d_objective_f32"] = r32.d_objective.mean(0).tolist()
        log(f"tangent {name} float32 fast path: {t32 * 1e3:.2f} ms")
    # 65..128 features: the float32 tangent (no float64 one holds G^-1 and its
    # tangent at this width) against one float64 primal transition
    X, y, _ = make_logistic(N_ROWS, 128, DATA_SEED)
    model, C, D = gj.LogisticRegression(X, y, PRIOR, device=dev), 256, 128
    integ = functools.partial(implicit_midpoint, max_iters=MAXIT_LOG)
    q = 0.5 * torch.randn(C, D, generator=torch.Generator().manual_seed(7), dtype=torch.float64)
    s64 = gj.rmhmc.init(q.to(dev), model)
    s32 = gj.rmhmc.init(q.to(dev, torch.float32), model)
    kern = build_kernel(integ)
    esjd_acceptance(gr.PRNGKey(1, 0), s32, model, EPS_LOG, gj.FisherMetric(), 32, integrator=integ)
    kern(gr.PRNGKey(1, 0), s64, model, EPS_LOG, gj.FisherMetric(), 32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(repeats):
        r32 = esjd_acceptance(gr.PRNGKey(2, i), s32, model, EPS_LOG, gj.FisherMetric(), 32, integrator=integ)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / repeats
    t0 = time.perf_counter()
    for i in range(repeats):
        kern(gr.PRNGKey(2, i), s64, model, EPS_LOG, gj.FisherMetric(), 32)
    torch.cuda.synchronize()
    t_pri = (time.perf_counter() - t0) / repeats
    out["logistic_2048x128_256_chains"] = {"tangent_f32_ms": t32 * 1e3, "primal_f64_transition_ms": t_pri * 1e3,
                                           "tangent_over_primal": t32 / t_pri, "num_integration_steps": 32,
                                           "mean_d_objective_f32": r32.d_objective.mean(0).tolist()}
    log(f"tangent logistic_2048x128_256_chains float32: {t32 * 1e3:.2f} ms vs primal {t_pri * 1e3:.2f} ms")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed transitions of the headline workload")
    ap.add_argument("--warmup", type=int, default=20, help="untimed transitions before timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 1024-chain and funnel comparisons")
    ap.add_argument("--only", choices=["headline", "tangent", "uneven4096", "funnel", "funnel_nuts", "logistic_nuts", "logistic_nuts_lmc", "lmc_alpha2", "logistic_lmc", "mmala",
                             "uneven", "dense_lmc", "dense_rmhmc"],
                    help="run one secondary workload alone and print its JSON (for profiling it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB three times (reference kernel for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the chains of this workload are "
              "independent, so more GPUs would each run their own share of them.", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        x = torch.ones(64 * 1024 * 1024, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {x.numel() * 4} bytes x 3")
        return
    K, W = max(1, args.steps), max(0, args.warmup)
    if args.only == "funnel":
        print(json.dumps(funnel_forms(K, W, dev, plain=not args.no_plain)), flush=True)
        return
    if args.only == "funnel_nuts":
        print(json.dumps(nuts_forms(K, W, dev)), flush=True)
        print(json.dumps(nuts_forms(K, W, dev, tuned=True)), flush=True)
        return
    if args.only == "logistic_nuts":
        print(json.dumps(logistic_nuts_forms(K, W, dev)), flush=True)
        print(json.dumps(logistic_nuts_forms(K, W, dev, tuned=True)), flush=True)
        return
    if args.only == "logistic_nuts_lmc":
        print(json.dumps(logistic_nuts_forms(K, W, dev, integrator="lmc")), flush=True)
        print(json.dumps(logistic_nuts_forms(K, W, dev, integrator="lmc", tuned=True)), flush=True)
        return
    if args.only == "uneven":
        print(json.dumps(uneven_batch_forms(32, K, W, dev, plain=not args.no_plain)), flush=True)
        return
    if args.only == "uneven4096":
        print(json.dumps(uneven_batch_forms(16, K, W, dev, plain=not args.no_plain)), flush=True)
        return
    if args.only == "headline":
        print(json.dumps(logistic_forms(CHAINS, K, W, dev, plain=not args.no_plain)), flush=True)
        return
    if args.only == "dense_rmhmc":
        print(json.dumps(dense_rmhmc_forms(dev)), flush=True)
        return
    if args.only == "dense_lmc":
        print(json.dumps(dense_lmc_forms(dev)), flush=True)
        return
    if args.only == "mmala":
        print(json.dumps(mmala_forms(K, W, dev)), flush=True)
        return
    if args.only == "tangent":
        print(json.dumps(tangent_forms(dev)), flush=True)
        return
    if args.only == "logistic_lmc":
        print(json.dumps(logistic_lmc_forms(K, W, dev)), flush=True)
        print(json.dumps(logistic_lmc_forms(K, W, dev, tuned=True)), flush=True)
        return
    if args.only == "lmc_alpha2":
        print(json.dumps(lmc_alpha2_forms(dev)), flush=True)
        return

    head = logistic_forms(CHAINS, K, W, dev, plain=not args.no_plain)
    tuned = head["tuned"]
    achieved = tuned["mfma_flops_per_launch"] / (tuned["avg_launch_ms"] / 1e3)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": BF16_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK, "traffic": measured_traffic(),
                "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                "kernel": f"logistic_rmhmc_f32_kernel<2,{tuned['waves_per_chain']},false>",
                "launches": tuned["launches"],
                "avg_launch_ms": tuned["avg_launch_ms"],
                "algorithmic_flops_per_launch": tuned["mfma_flops_per_launch"]}

    extra = {}
    if not args.no_extra:
        extra["logistic_1024_chains"] = logistic_forms(1024, K, W, dev, plain=not args.no_plain)
        extra["uneven_batch_8192_chains"] = uneven_batch_forms(32, K, W, dev, plain=not args.no_plain)
        extra["uneven_batch_4096_chains"] = uneven_batch_forms(16, K, W, dev, plain=not args.no_plain)
        extra["funnel_d100_8192_chains"] = funnel_forms(300, 100, dev, plain=not args.no_plain)
        extra["funnel_nuts_d100_8192_chains"] = nuts_forms(100, 10, dev,
                                                           extra["funnel_d100_8192_chains"].get("plain"))
        extra["funnel_nuts_d100_8192_chains_f32"] = nuts_forms(100, 10, dev,
                                                               extra["funnel_d100_8192_chains"].get("plain"),
                                                               tuned=True)
        extra["logistic_nuts_1024_chains"] = logistic_nuts_forms(10, 2, dev)
        extra["logistic_nuts_1024_chains_f32"] = logistic_nuts_forms(10, 2, dev, tuned=True)
        extra["logistic_nuts_1024_chains_f32"]["tuned_vs_f64"] = {
            k: extra["logistic_nuts_1024_chains_f32"][k] / extra["logistic_nuts_1024_chains"][k]
            for k in ("value", "min_ess_per_sec")}
        extra["logistic_lmc_1024_chains"] = logistic_lmc_forms(20, 5, dev)
        extra["logistic_lmc_1024_chains_f32"] = logistic_lmc_forms(20, 5, dev, tuned=True)
        extra["logistic_lmc_1024_chains_f32"]["tuned_vs_f64"] = {
            k: extra["logistic_lmc_1024_chains_f32"][k] / extra["logistic_lmc_1024_chains"][k]
            for k in ("value", "min_ess_per_sec")}
        extra["logistic_nuts_lmc_1024_chains"] = logistic_nuts_forms(10, 2, dev, integrator="lmc")
        extra["logistic_nuts_lmc_1024_chains_f32"] = logistic_nuts_forms(10, 2, dev, integrator="lmc", tuned=True)
        extra["logistic_nuts_lmc_1024_chains_f32"]["tuned_vs_f64"] = {
            k: extra["logistic_nuts_lmc_1024_chains_f32"][k] / extra["logistic_nuts_lmc_1024_chains"][k]
            for k in ("value", "min_ess_per_sec")}
        extra["mmala_1024_chains"] = mmala_forms(200, 100, dev)
        extra["dense_lmc_funnel_d20_1024_chains"] = dense_lmc_forms(dev)
        extra["dense_rmhmc_funnel_d20_1024_chains"] = dense_rmhmc_forms(dev)
        extra["funnel_lmcmonge_d100_8192_chains"] = funnel_lmc_forms(300, 100, dev, plain=not args.no_plain)
        extra["tangent"] = tangent_forms(dev)
        extra["lmc_alpha2_tangent"] = lmc_alpha2_forms(dev)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle restatement) ...")
        cpu = cpu_baseline()
        log(f"cpu: {cpu['value']:.3e} leapfrog-steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": tuned["value"],
        "unit": "leapfrog-steps/s",
        "n_gpus": 1,
        "steps": tuned["steps"],
        "warmup": W,
        "ms_per_step": tuned["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Bayesian logistic regression RMHMC, Fisher metric, X 2048x64 iid N(0,1/64), "
                               "prior N(0,100 I), data seed 2",
                   "chains": CHAINS, "dim": D_LOG, "rows": N_ROWS, "step_size": EPS_LOG,
                   "num_integration_steps": L_LOG, "max_iters": MAXIT_LOG, "init": "N(0, I) seed 1",
                   "transitions_per_launch": CHUNK_LOG},
        "min_ess_per_sec": tuned["min_ess_per_sec"],
        "min_ess": tuned["min_ess"],
        "ess_window": {k: head["ess_window"][k] for k in ("window_transitions", "draws")}
        | {"warmup_transitions": max(WARM_MIN, W) // CHUNK_LOG * CHUNK_LOG},
        "acceptance_rate": tuned["acceptance_rate"],
        "plain": head["plain"],
        "tuned_vs_plain": head.get("tuned_vs_plain"),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
