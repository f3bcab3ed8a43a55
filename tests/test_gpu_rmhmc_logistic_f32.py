"""GPU: the tuned RMHMC Fisher-logistic path (float32 state, bf16 MFMA
contractions with split operands, float32 sweep + refinement, float64 energies)
against float64 on the same q0, z, u.

The reference is the float64 oracle (dense Fisher metric, Cholesky inside the
Gaussian log-density, autodiff) up to the 2048 x 64 configuration and the
128-feature capacity; for more chains, several transitions and the
multi-posterior batches it is the float64 HIP kernel, itself pinned to the
oracle at 1e-10.  Stated tolerances of the tuned path:
  * end state (new position, proposal position): 1e-5 norm-wise relative;
  * momentum p0 = chol(G) z: 1e-5; velocity and gradient at the proposal: 1e-4;
  * energy and energy difference dH: 1e-3 absolute;
  * accept decisions identical except where |log u - dH| < 1e-3.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import build_kernel, implicit_midpoint
from geomjax_amd.synthetic import make_logistic
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL_STATE, TOL_MOM, TOL_VEL, TOL_H = 1e-5, 1e-5, 1e-4, 1e-3


def _check_step(info, s, ref_info, ref_state, u, q_prev, label):
    """Compare a float32 step (info, s) with a float64 step (ref_info, ref_state)."""
    w64 = ref_info.proposal.weight.double().cpu()
    ties = (torch.log(u.double().cpu()) - w64).abs() < TOL_H
    keep = ~ties
    errs = {
        "momentum": rel_err(info.momentum.double(), ref_info.momentum.double()),
        "proposal.position": rel_err(info.proposal.state.position.double(), ref_info.proposal.state.position.double()),
        "proposal.velocity": rel_err(info.proposal.state.velocity.double(), ref_info.proposal.state.velocity.double()),
        "proposal.logdensity_grad": rel_err(info.proposal.state.logdensity_grad.double(),
                                            ref_info.proposal.state.logdensity_grad.double()),
        "energy": (info.energy.cpu() - ref_info.energy.cpu()).abs().max().item(),
        "dH": (info.proposal.weight.cpu() - w64).abs().max().item(),
        "position": rel_err(s.position.double().cpu()[keep], ref_state.position.double().cpu()[keep]),
    }
    print(f"  {label}: ties {int(ties.sum())}, accepted {int(info.is_accepted.sum())}, "
          + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert torch.equal(info.is_accepted.cpu()[keep], ref_info.is_accepted.cpu()[keep])
    assert errs["momentum"] < TOL_MOM
    assert errs["proposal.position"] < TOL_STATE
    assert errs["position"] < TOL_STATE
    assert errs["proposal.velocity"] < TOL_VEL
    assert errs["proposal.logdensity_grad"] < TOL_VEL
    assert errs["energy"] < TOL_H
    assert errs["dH"] < TOL_H


@pytest.mark.parametrize("N,D,C,eps,L,max_iters,seed", [
    (40, 6, 3, 0.1, 6, 6, 5),      # one 64-row group: the smallest row loops
    (256, 8, 3, 0.1, 10, 6, 0),
    (300, 13, 2, 0.1, 6, 6, 1),    # ragged: N not a multiple of 32
    (1000, 40, 2, 0.1, 5, 6, 3),   # two 32-feature blocks with padding
    (2048, 64, 2, 0.1, 10, 6, 2),  # the 2048 x 64 configuration itself
    (1024, 128, 2, 0.05, 6, 6, 4), # the four-wave kernel at its capacity
])
def test_step_f32_matches_oracle(N, D, C, eps, L, max_iters, seed):
    print(f"case N={N} D={D} C={C} eps={eps} L={L} max_iters={max_iters}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    M = om.LogisticRegression(X, y, 100.0)
    q0 = seeded_normal(seed, C, D)
    z, u = seeded_normal(500 + seed, C, D), seeded_uniform(600 + seed, C)
    so = orm.init(q0, M.logp)
    so1, io = orm.build_kernel(max_iters=max_iters)((z, u), so, M.logp, eps, M.fisher_metric_fn, L)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.rmhmc.init(q0.float().to(DEV), model)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=max_iters))
    s1, info = kern(gr.HostNoise(z.float().to(DEV), u.to(DEV)), s, model, eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    assert s1.position.dtype == torch.float32
    _check_step(info, s1, io, so1, u, q0, "f32 vs oracle")


def test_step_f32_matches_f64_kernel_2048x64():
    N, D, C, eps, L, max_iters = 2048, 64, 16, 0.1, 10, 6
    X, y, _ = make_logistic(N, D, 2)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = seeded_normal(11, C, D).to(DEV)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters)
    kern = build_kernel(integ)
    s64 = gj.rmhmc.init(q0, model)
    s32 = gj.rmhmc.init(q0.float(), model)
    for t in range(2):
        print(f"case N={N} D={D} C={C} transition {t}", flush=True)
        z, u = seeded_normal(700 + t, C, D).to(DEV), seeded_uniform(800 + t, C).to(DEV)
        s64n, i64 = kern(gr.HostNoise(z, u), s64, model, eps, gj.FisherMetric(), L)
        s32n, i32 = kern(gr.HostNoise(z.float(), u), s32, model, eps, gj.FisherMetric(), L)
        torch.cuda.synchronize()
        _check_step(i32, s32n, i64, s64n, u, q0, f"2048x64 t={t}")
        # continue both from the float64 state so the comparison stays one transition deep
        s64 = s64n
        s32 = gj.rmhmc.init(s64n.position.float(), model)


@pytest.mark.parametrize("N,D,C,seed", [
    (500, 100, 4, 5),    # four 32-feature blocks with padding
    (1024, 128, 4, 6),   # at the capacity of the four-wave kernel
])
def test_step_f32_wide_matches_f64_kernel(N, D, C, seed):
    eps, L, max_iters = 0.05, 6, 6
    X, y, _ = make_logistic(N, D, seed)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = (0.5 * seeded_normal(30 + seed, C, D)).to(DEV)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=max_iters))
    s64 = gj.rmhmc.init(q0, model)
    s32 = gj.rmhmc.init(q0.float(), model)
    for t in range(2):
        print(f"case N={N} D={D} C={C} transition {t}", flush=True)
        z, u = seeded_normal(40 + 10 * seed + t, C, D).to(DEV), seeded_uniform(60 + 10 * seed + t, C).to(DEV)
        s64n, i64 = kern(gr.HostNoise(z, u), s64, model, eps, gj.FisherMetric(), L)
        s32n, i32 = kern(gr.HostNoise(z.float(), u), s32, model, eps, gj.FisherMetric(), L)
        torch.cuda.synchronize()
        _check_step(i32, s32n, i64, s64n, u, q0, f"{N}x{D} t={t}")
        s64 = s64n
        s32 = gj.rmhmc.init(s64n.position.float(), model)


@pytest.mark.parametrize("shapes", [
    [(200, 8), (700, 24), (1500, 40), (2048, 64), (333, 32)],
    [(200, 8), (600, 128), (1500, 40), (400, 72), (333, 32), (2048, 100)],  # all three packed widths
])
def test_batch_f32_matches_f64_kernel(shapes):
    models = [gj.LogisticRegression(*make_logistic(n, d, 20 + i)[:2], 100.0) for i, (n, d) in enumerate(shapes)]
    batch = gj.LogisticBatch(models, 3)
    C, Dm, eps, L = batch.num_chains, batch.dim, 0.08, 6
    print(f"case batch {shapes} x 3 chains", flush=True)
    q0 = torch.zeros(C, Dm, dtype=torch.float64)
    for c in range(C):
        d = shapes[c // 3][1]
        q0[c, :d] = seeded_normal(900 + c, d)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    for c in range(C):
        d = shapes[c // 3][1]
        z[c, :d] = seeded_normal(950 + c, d)
    u = seeded_uniform(990, C)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=6))
    s64 = gj.rmhmc.init(q0.to(DEV), batch)
    s32 = gj.rmhmc.init(q0.float().to(DEV), batch)
    s64n, i64 = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s64, batch, eps, gj.FisherMetric(), L)
    s32n, i32 = kern(gr.HostNoise(z.float().to(DEV), u.to(DEV)), s32, batch, eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    _check_step(i32, s32n, i64, s64n, u, q0, "batch")
    # coordinates past a chain's own D stay zero
    for c in range(C):
        d = shapes[c // 3][1]
        assert torch.all(s32n.position[c, d:] == 0)


@pytest.mark.parametrize("N,D", [(512, 40), (384, 100)])
def test_sample_kernel_matches_step_kernel_f32(N, D):
    from geomjax_amd.sampling import run_chains

    C, eps, L, T = 24, 0.1, 5, 3
    X, y, _ = make_logistic(N, D, 7)
    model = gj.LogisticRegression(X, y, 100.0)
    alg = gj.rmhmc(model, eps, gj.FisherMetric(), L, integrator=functools.partial(implicit_midpoint, max_iters=6))
    q0 = seeded_normal(71, C, D).float().to(DEV)
    zs, us = seeded_normal(72, T, C, D).float().to(DEV), seeded_uniform(73, T, C).to(DEV)
    s = alg.init(q0)
    res = run_chains(alg, s, T, gr.HostNoise(zs, us))
    for t in range(T):
        s, _ = alg.step(gr.HostNoise(zs[t], us[t]), s)
        assert rel_err(res.samples[t], s.position) <= 1e-6
    torch.cuda.synchronize()
    assert rel_err(res.state.position, s.position) <= 1e-6
    assert rel_err(res.state.logdensity, s.logdensity) <= 1e-9
    assert rel_err(res.state.logdensity_grad, s.logdensity_grad) <= 1e-5
    assert int(res.num_fevals.min()) >= T * L * 2


def test_device_prng_sampling_f32():
    N, D, C, eps, L, T = 2048, 64, 64, 0.1, 10, 4
    X, y, _ = make_logistic(N, D, 2)
    model = gj.LogisticRegression(X, y, 100.0)
    alg = gj.rmhmc(model, eps, gj.FisherMetric(), L, integrator=functools.partial(implicit_midpoint, max_iters=6))
    s = alg.init(seeded_normal(5, C, D).float().to(DEV))
    res = alg.sample_fn(gr.key(3), s, T)
    torch.cuda.synchronize()
    acc = res.acceptance_rate.mean().item()
    print(f"  acceptance {acc:.3f}, fevals/chain/transition {res.num_fevals.double().mean().item() / T:.1f}")
    assert torch.isfinite(res.samples).all()
    assert 0.5 < acc <= 1.0
    assert int(res.num_fevals.max()) <= T * L * 8


def _layouts_numpy(X, y):
    """The two fragment layouts and labels of one posterior, restated with numpy."""
    import numpy as np

    N, D = X.shape
    npad, nb = -(-N // 64) * 64, 1 if D <= 32 else (2 if D <= 64 else 4)
    Xp = np.zeros((npad, 32 * nb))
    Xp[:N, :D] = X
    lane = np.arange(64)
    xg = np.empty((npad // 16, nb, 64, 8))
    for kb in range(npad // 16):
        for b in range(nb):
            for i in range(8):
                xg[kb, b, :, i] = Xp[kb * 16 + (lane // 32) * 8 + i, b * 32 + lane % 32]
    xr = np.empty((npad // 32, 2 * nb, 64, 8))
    for rb in range(npad // 32):
        for kc in range(2 * nb):
            for i in range(8):
                xr[rb, kc, :, i] = Xp[rb * 32 + lane % 32, kc * 16 + (lane // 32) * 8 + i]
    yf = np.zeros(npad)
    yf[:N] = y
    return xg.reshape(-1), xr.reshape(-1), yf


def test_pack_layout_matches_numpy_restatement():
    import numpy as np

    shapes = [(70, 5), (64, 32), (333, 47), (100, 90)]
    data = [make_logistic(n, d, 40 + i)[:2] for i, (n, d) in enumerate(shapes)]
    batch = gj.LogisticBatch([gj.LogisticRegression(X, y, 100.0) for X, y in data], 1)
    buf = batch._table.packed().cpu().numpy()
    hdr = np.frombuffer(buf[:40 * len(shapes)].tobytes(), dtype=np.dtype(
        [("xg", "<i8"), ("xr", "<i8"), ("xs", "<i8"), ("yf", "<i8"), ("npad", "<i4"), ("nb", "<i4")]))
    flt = np.frombuffer(buf.tobytes(), dtype=np.float32)
    u16 = np.frombuffer(buf.tobytes(), dtype=np.uint16)
    for p, (X, y) in enumerate(data):
        xg, xr, yf = _layouts_numpy(X, y)
        h = hdr[p]
        d = X.shape[1]
        assert h["npad"] == -(-X.shape[0] // 64) * 64 and h["nb"] == (1 if d <= 32 else (2 if d <= 64 else 4))
        np.testing.assert_array_equal(flt[h["xg"]:h["xg"] + xg.size], xg.astype(np.float32))
        np.testing.assert_array_equal(flt[h["xr"]:h["xr"] + xr.size], xr.astype(np.float32))
        np.testing.assert_array_equal(flt[h["yf"]:h["yf"] + yf.size], yf.astype(np.float32))
        # split layout: per lane fragment 8 bf16 hi then 8 bf16 lo with hi + lo ~ x (|x - hi - lo| <= 2^-17 |x|)
        pairs = u16[2 * h["xs"]:2 * h["xs"] + 2 * xr.size].reshape(-1, 2, 8)
        bf = lambda b: (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        hi, lo = bf(pairs[:, 0, :]).reshape(-1), bf(pairs[:, 1, :]).reshape(-1)
        x32 = xr.astype(np.float32).astype(np.float64)
        assert np.all(np.abs(x32 - hi - lo) <= 2.0 ** -17 * np.abs(x32))
        assert np.all(np.abs(lo) <= 2.0 ** -8 * np.abs(hi) + 1e-45)


def test_uneven_batch_full_size_matches_each_posterior_and_f64():
    """The uneven-batch configuration at full size: 256 posteriors (D in
    {8, 16, 32, 64, 128}, N up to 4096), cost-ordered dispatch and concurrent
    per-width launches, two chains per posterior.  The float64 batch equals
    every posterior run alone (1e-12); the tuned batch equals the float64 batch
    at the tuned bars."""
    from geomjax_amd.synthetic import make_uneven_batch

    data = make_uneven_batch(P=256, seed=3)
    cpp, eps, L = 2, 0.1, 10
    models = [gj.LogisticRegression(X, y, 100.0) for X, y, _ in data]
    batch = gj.LogisticBatch(models, cpp)
    C, Dm = batch.num_chains, batch.dim
    dims = [m.D for m in models]
    print(f"case uneven batch: {len(models)} posteriors, N max {batch.N}, {C} chains", flush=True)
    q0 = torch.zeros(C, Dm, dtype=torch.float64)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    for c in range(C):
        d = dims[c // cpp]
        q0[c, :d] = 0.5 * seeded_normal(3000 + c, d)
        z[c, :d] = seeded_normal(4000 + c, d)
    u = seeded_uniform(5000, C)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=6))
    s64 = gj.rmhmc.init(q0.to(DEV), batch)
    s32 = gj.rmhmc.init(q0.float().to(DEV), batch)
    s64n, i64 = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s64, batch, eps, gj.FisherMetric(), L)
    s32n, i32 = kern(gr.HostNoise(z.float().to(DEV), u.to(DEV)), s32, batch, eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    rows_err = []
    for p, m in enumerate(models):
        rows = slice(p * cpp, (p + 1) * cpp)
        e = rel_err(i32.proposal.state.position[rows, :m.D], i64.proposal.state.position[rows, :m.D])
        ev = rel_err(i32.proposal.state.velocity[rows, :m.D], i64.proposal.state.velocity[rows, :m.D])
        rows_err.append((e, ev, p, m.N, m.D))
    rows_err.sort(reverse=True)
    print("  worst posteriors (proposal position err, velocity err, index, N, D):",
          [(f"{a:.2e}", f"{b:.2e}", p, n, d) for a, b, p, n, d in rows_err[:8]], flush=True)
    _check_step(i32, s32n, i64, s64n, u, q0, "uneven batch, tuned vs float64")
    worst = 0.0
    for p, m in enumerate(models):
        rows = slice(p * cpp, (p + 1) * cpp)
        d = m.D
        sa = gj.rmhmc.init(q0[rows, :d].contiguous().to(DEV), m)
        sa1, ia = kern(gr.HostNoise(z[rows, :d].contiguous().to(DEV), u[rows].to(DEV)), sa, m, eps,
                       gj.FisherMetric(), L)
        for a, b in ((s64n.position[rows, :d], sa1.position), (i64.momentum[rows, :d], ia.momentum),
                     (i64.proposal.state.position[rows, :d], ia.proposal.state.position),
                     (i64.energy[rows], ia.energy)):
            e = rel_err(a, b)
            worst = max(worst, e)
            assert e < 1e-12, f"posterior {p} ({m.N}x{d}): batch differs from the posterior alone by {e:.2e}"
        assert torch.equal(i64.is_accepted[rows].cpu(), ia.is_accepted.cpu())
        assert torch.all(s64n.position[rows, d:] == 0) and torch.all(s32n.position[rows, d:] == 0)
    torch.cuda.synchronize()
    print(f"  float64 batch vs each posterior alone: worst {worst:.2e}", flush=True)


@pytest.mark.parametrize("waves", [1, 2, (2, 1), (1, 2)])
def test_every_workgroup_shape_matches_f64_kernel(waves):
    """The tuned kernels' launch variants (one or two waves per chain, per
    packed width) all equal the float64 kernel at the tuned bars, alone and in
    a batch of every packed width."""
    eps, L = 0.1, 8
    kern64 = build_kernel(functools.partial(implicit_midpoint, max_iters=6))
    kern32 = build_kernel(functools.partial(implicit_midpoint, max_iters=6), waves_per_chain=waves)
    for (N, D, C) in [(256, 8, 6), (64, 20, 4), (130, 33, 4), (2048, 64, 8)]:
        X, y, _ = make_logistic(N, D, 60 + D)
        model = gj.LogisticRegression(X, y, 100.0)
        q0 = (0.5 * seeded_normal(61 + D, C, D)).to(DEV)
        z, u = seeded_normal(62 + D, C, D).to(DEV), seeded_uniform(63 + D, C).to(DEV)
        print(f"case waves={waves} N={N} D={D} C={C}", flush=True)
        s64n, i64 = kern64(gr.HostNoise(z, u), gj.rmhmc.init(q0, model), model, eps, gj.FisherMetric(), L)
        s32n, i32 = kern32(gr.HostNoise(z.float(), u), gj.rmhmc.init(q0.float(), model), model, eps,
                           gj.FisherMetric(), L)
        torch.cuda.synchronize()
        _check_step(i32, s32n, i64, s64n, u, q0, f"{N}x{D} waves={waves}")
    shapes = [(200, 8), (600, 128), (1500, 40), (333, 32), (2048, 64)]
    models = [gj.LogisticRegression(*make_logistic(n, d, 70 + i)[:2], 100.0) for i, (n, d) in enumerate(shapes)]
    batch = gj.LogisticBatch(models, 3)
    C, Dm = batch.num_chains, batch.dim
    q0 = torch.zeros(C, Dm, dtype=torch.float64)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    for c in range(C):
        d = shapes[c // 3][1]
        q0[c, :d] = 0.5 * seeded_normal(1100 + c, d)
        z[c, :d] = seeded_normal(1200 + c, d)
    u = seeded_uniform(1300, C)
    print(f"case waves={waves} batch {shapes} x 3 chains", flush=True)
    s64n, i64 = kern64(gr.HostNoise(z.to(DEV), u.to(DEV)), gj.rmhmc.init(q0.to(DEV), batch), batch, eps,
                       gj.FisherMetric(), L)
    s32n, i32 = kern32(gr.HostNoise(z.float().to(DEV), u.to(DEV)), gj.rmhmc.init(q0.float().to(DEV), batch), batch,
                       eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    _check_step(i32, s32n, i64, s64n, u, q0, f"batch waves={waves}")


def test_autotuner_choice_is_cached_per_problem():
    from geomjax_amd import autotune

    X, y, _ = make_logistic(512, 40, 80)
    model = gj.LogisticRegression(X, y, 100.0)
    field, timings = autotune.tune_logistic(model, 256)
    print(f"  chosen waves_per_chain field {field:#x}, timings {timings}", flush=True)
    assert field in autotune.LOGISTIC_WAVES
    assert len(timings) == 2  # one width present: two distinct candidates
    assert autotune.logistic_waves_field(model, 256) == field
    assert autotune.tune_logistic(model, 256) == (field, {})
