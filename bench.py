#!/usr/bin/env python
"""Benchmark of the Riemannian integrator hot path on one MI355X.

Workload: Neal's funnel D=100 (1 log-scale + 99 conditionals), RMHMC with the
Monge metric G = I + alpha2 grad l grad l^T (alpha2 = 1), step 0.03, 32
implicit-midpoint steps per transition, fixed point capped at 8 iterations,
8192 chains, initial positions N(0, I) with seed 1.

A "step" is one transition of all 8192 chains (32 integrator steps each).

* tuned form (the reported ``value``): FP32 state with FP64 energy /
  convergence test, autotuned lanes-per-chain, persistent launches of 100
  transitions each with the chain state held in registers, device Philox
  noise.
* plain form: FP64 kernels, library-default launch shape, one launch per
  transition from the host.

``value`` = chains * 32 * steps / wall seconds of the timed region (inputs
already resident in HBM).  min-ESS/s uses the multi-chain ESS of
geomjax/diagnostics.py:78-209 over the timed samples.  The CPU baseline is the
float64 oracle restatement (dense metric + autodiff as the reference) timed on
this host's cores on a bounded sample.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "min-ESS/sec & leapfrog-steps/sec, 4096 chains, tuned vs plain GPU form"
D, CHAINS, EPS, L, MAX_ITERS, ALPHA2, SIGMA, SEED = 100, 8192, 0.03, 32, 8, 1.0, 3.0, 1
CHUNK = 100  # transitions per persistent launch
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_rmhmc_monge_funnel(num_fevals_total, chains, transitions):
    """Algorithmic flops of RMHMC-Monge on the funnel (transcendentals count 1).

    per evaluation of the vector field 33 D + 25 flops + 1 exp; per chain-step
    11 D; per transition two energies (8 D + 2 log) and the rank-one momentum
    draw (10 D).
    """
    return (num_fevals_total * (33 * D + 26) + chains * transitions * L * 11 * D
            + chains * transitions * (18 * D + 2))


def min_ess(samples_list):
    """min over coordinates of the multi-chain ESS of [T, C, D] samples (chunked over D)."""
    import torch

    from geomjax_amd.diagnostics import effective_sample_size

    x = torch.cat(samples_list, dim=0) if isinstance(samples_list, list) else samples_list
    out = []
    for d0 in range(0, x.shape[2], 10):
        out.append(effective_sample_size(x[:, :, d0:d0 + 10], chain_axis=1, sample_axis=0))
    ess = torch.cat([o.reshape(-1) for o in out])
    return float(ess.min().item()), int(ess.argmin().item())


def cpu_baseline(seconds_budget=30.0):
    """Oracle (float64 dense + autodiff restatement) on this host's cores."""
    import torch

    from oracle import models as om
    from oracle import rmhmc as orm

    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    M = om.NealFunnel(D, sigma=SIGMA)
    kern = orm.build_kernel(max_iters=MAX_ITERS)
    metric = om.monge_metric_fn(M.logp, ALPHA2, D)
    g = torch.Generator().manual_seed(SEED)
    C = 64
    q0 = torch.randn(C, D, generator=g, dtype=torch.float64)
    s = orm.init(q0, M.logp)
    # untimed warm-up on a few chains
    kern((torch.randn(4, D, generator=g, dtype=torch.float64), torch.rand(4, generator=g, dtype=torch.float64)),
         orm.init(q0[:4], M.logp), M.logp, EPS, metric, L)
    n_tr, t_used = 0, 0.0
    while True:
        z = torch.randn(C, D, generator=g, dtype=torch.float64)
        u = torch.rand(C, generator=g, dtype=torch.float64)
        t0 = time.perf_counter()
        s, _ = kern((z, u), s, M.logp, EPS, metric, L)
        t_used += time.perf_counter() - t0
        n_tr += 1
        if t_used > seconds_budget / 2 or n_tr >= 4:
            break
    return {"value": C * L * n_tr / t_used, "unit": "leapfrog-steps/s", "cores": threads, "kind": "port",
            "sample": f"{C} chains x {n_tr} transitions of the same workload (float64 dense-metric autodiff "
                      f"restatement, {t_used:.1f} s)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed transitions (multiple of 100)")
    ap.add_argument("--warmup", type=int, default=100, help="untimed burn-in transitions")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the chains of this workload are "
              "independent, so more GPUs would each run their own share; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    import torch

    import geomjax_amd as gj
    from geomjax_amd import autotune
    from geomjax_amd import random as gr
    from geomjax_amd.rmhmc import build_kernel, implicit_midpoint, sample

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    K = max(CHUNK, (args.steps // CHUNK) * CHUNK)
    W = max(0, args.warmup)
    model, metric = gj.NealFunnel(D, sigma=SIGMA), gj.MongeMetric(ALPHA2)
    integ = __import__("functools").partial(implicit_midpoint, max_iters=MAX_ITERS)
    q_init = torch.randn(CHAINS, D, generator=torch.Generator().manual_seed(SEED), dtype=torch.float64)

    # ------------------------------------------------------------------ tuned
    s32 = gj.rmhmc.init(q_init.to(dev, torch.float32), model)
    key = gr.key(SEED)

    def run(lpc, state=s32, k=key, T=CHUNK, out=None):
        return sample(k, state, model, EPS, metric, L, T, integrator=integ, lanes_per_chain=lpc,
                      store_samples=out is not None, samples_out=out)

    lpc, timings = autotune.tune("rmhmc_monge_funnel", D, CHAINS, torch.float32, lambda l: run(l), repeats=2)
    log(f"autotune lanes_per_chain: {lpc}  ({ {k: round(v * 1e3, 2) for k, v in timings.items()} } ms / {CHUNK} tr)")
    counter = 10_000
    state = s32
    for _ in range(max(1, W // CHUNK)):
        r = run(lpc, state, gr.PRNGKey(key.seed, counter))
        state, counter = r.state, counter + CHUNK
    samples = torch.empty(K, CHAINS, D, dtype=torch.float32, device=dev)
    fev_list, accs = [], []
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for c in range(K // CHUNK):
        r = run(lpc, state, gr.PRNGKey(key.seed, counter), out=samples[c * CHUNK:(c + 1) * CHUNK])
        state, counter = r.state, counter + CHUNK
        fev_list.append(r.num_fevals)
        accs.append(r.acceptance_rate)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = ev0.elapsed_time(ev1) / 1e3
    n_launch = K // CHUNK
    value = CHAINS * L * K / wall
    acc_mean = float(torch.stack(accs).mean().item())
    fev_total = int(torch.stack(fev_list).sum().item())
    ess_min, ess_arg = min_ess(samples)
    log(f"tuned: {K} transitions in {wall:.3f} s ({value:.3e} leapfrog-steps/s), kernel {kernel_s:.3f} s, "
        f"acc {acc_mean:.3f}, fevals/chain/transition {fev_total / CHAINS / K:.2f}, min-ESS {ess_min:.0f} "
        f"(coord {ess_arg})")
    flops = flops_rmhmc_monge_funnel(fev_total, CHAINS, K)
    achieved = flops / n_launch / (kernel_s / n_launch)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK, "traffic": None,
                "kernel": f"funnel_rmhmc_kernel<float,{lpc},*,false>", "launches": n_launch,
                "avg_launch_ms": kernel_s / n_launch * 1e3,
                "algorithmic_flops_per_launch": flops / n_launch}
    del samples

    # ------------------------------------------------------------------ plain
    plain = None
    if not args.no_plain:
        s64 = gj.rmhmc.init(q_init.to(dev), model)
        kern = build_kernel(integ)
        for t in range(min(W, 20)):
            s64, _ = kern(gr.PRNGKey(key.seed + 1, t), s64, model, EPS, metric, L)
        psamples = torch.empty(K, CHAINS, D, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(K):
            s64, info = kern(gr.PRNGKey(key.seed + 1, 1000 + t), s64, model, EPS, metric, L)
            psamples[t].copy_(s64.position)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pess, _ = min_ess(psamples)
        del psamples
        plain = {"value": CHAINS * L * K / pwall, "unit": "leapfrog-steps/s", "min_ess_per_sec": pess / pwall,
                 "min_ess": pess, "dtype": "f64", "form": "FP64 kernels, default launch shape, one launch per transition"}
        log(f"plain: {plain['value']:.3e} leapfrog-steps/s, min-ESS/s {plain['min_ess_per_sec']:.3e}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle restatement) ...")
        cpu = cpu_baseline()
        log(f"cpu: {cpu['value']:.3e} leapfrog-steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "leapfrog-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Neal's funnel D=100, RMHMC implicit midpoint, Monge metric alpha2=1",
                   "chains": CHAINS, "dim": D, "step_size": EPS, "num_integration_steps": L,
                   "max_iters": MAX_ITERS, "init": "N(0, I) seed 1", "transitions_per_launch": CHUNK,
                   "lanes_per_chain": lpc},
        "min_ess_per_sec": ess_min / wall,
        "min_ess": ess_min,
        "acceptance_rate": acc_mean,
        "plain": plain,
        "tuned_vs_plain": None if plain is None else {
            "leapfrog_steps_per_sec": value / plain["value"],
            "min_ess_per_sec": (ess_min / wall) / plain["min_ess_per_sec"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
