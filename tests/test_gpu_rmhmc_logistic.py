"""GPU parity: RMHMC with the Fisher metric on Bayesian logistic regression (float64).

HIP kernel (MFMA Gram and quadratic forms, sweep inverse) through the C ABI
vs the float64 oracle (dense Fisher metric, Cholesky inside the Gaussian
log-density, autodiff) on the same q0, z, u.  Tolerances as the funnel tests:
1e-10 norm-wise relative on every field, dH within 1e-10 of the energy scale,
identical fixed-point iteration counts, equal accept decisions except
|log u - dH| < 1e-8.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import RMHMCState, build_kernel, implicit_midpoint
from geomjax_amd.synthetic import make_logistic
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("N,D,C,eps,L,max_iters,T,seed", [
    (256, 8, 3, 0.1, 10, 6, 2, 0),
    (300, 13, 2, 0.1, 6, 6, 1, 1),     # ragged: N not a multiple of 32, D not of 16
    (2048, 64, 2, 0.1, 10, 6, 1, 2),   # the 2048 x 64 logistic configuration's shape
    (500, 100, 2, 0.05, 4, 6, 1, 3),   # wide tiling (65..128 features), ragged D
    (1024, 128, 2, 0.05, 4, 6, 1, 4),  # wide tiling at its capacity
])
def test_step_f64_matches_oracle(N, D, C, eps, L, max_iters, T, seed):
    print(f"case N={N} D={D} C={C} eps={eps} L={L} max_iters={max_iters} T={T}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    M = om.LogisticRegression(X, y, 100.0)
    q0 = seeded_normal(seed, C, D)
    zs = [seeded_normal(500 + seed * 10 + t, C, D) for t in range(T)]
    us = [seeded_uniform(600 + seed * 10 + t, C) for t in range(T)]
    okern = orm.build_kernel(max_iters=max_iters)
    so = orm.init(q0, M.logp)
    model = gj.LogisticRegression(X, y, 100.0)
    metric = gj.FisherMetric()
    s = gj.rmhmc.init(q0.to(DEV), model)
    assert rel_err(s.logdensity, so.logdensity) < 1e-12
    assert rel_err(s.logdensity_grad, so.logdensity_grad) < 1e-12
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=max_iters), return_fixed_point_info=True)
    for t in range(T):
        tr = []
        so, io = okern((zs[t], us[t]), so, M.logp, eps, M.fisher_metric_fn, L, trace=tr)
        s, info, fp = kern(gr.HostNoise(zs[t].to(DEV), us[t].to(DEV)), s, model, eps, metric, L)
        torch.cuda.synchronize()
        ties = (torch.log(us[t]) - io.proposal.weight).abs() < 1e-8
        print(f"  t={t}: accepted {int(io.is_accepted.sum())}/{C}, ties {int(ties.sum())}, "
              f"iters {[x.iters.tolist() for x in tr][:3]}", flush=True)
        assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])
        assert rel_err(info.momentum, io.momentum) < 1e-10
        hscale = max(io.energy.abs().max().item(), 1.0)
        assert (info.energy.cpu() - io.energy).abs().max().item() < 1e-10 * hscale
        assert (info.proposal.weight.cpu() - io.proposal.weight).abs().max().item() < 1e-10 * hscale
        assert rel_err(info.acceptance_rate, io.acceptance_rate) < 1e-8
        ps, pso = info.proposal.state, io.proposal.state
        for name in ("position", "momentum", "velocity", "logdensity", "logdensity_grad"):
            e = rel_err(getattr(ps, name), getattr(pso, name))
            assert e < 1e-10, (name, e)
        assert torch.equal(fp.iters.cpu().long(), torch.stack([x.iters for x in tr], dim=1))
        keep = ~ties
        assert rel_err(s.position.cpu()[keep], so.position[keep]) < 1e-10
        assert rel_err(s.logdensity.cpu()[keep], so.logdensity[keep]) < 1e-10
        s = RMHMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV))


def test_sample_kernel_matches_step_kernel():
    from geomjax_amd.sampling import run_chains

    N, D, C, eps, L, T = 512, 16, 24, 0.1, 5, 3
    X, y, _ = make_logistic(N, D, 7)
    model = gj.LogisticRegression(X, y, 100.0)
    alg = gj.rmhmc(model, eps, gj.FisherMetric(), L, integrator=functools.partial(implicit_midpoint, max_iters=6))
    q0 = seeded_normal(71, C, D).to(DEV)
    zs, us = seeded_normal(72, T, C, D).to(DEV), seeded_uniform(73, T, C).to(DEV)
    s = alg.init(q0)
    res = run_chains(alg, s, T, gr.HostNoise(zs, us))
    for t in range(T):
        s, _ = alg.step(gr.HostNoise(zs[t], us[t]), s)
        assert rel_err(res.samples[t], s.position) <= 1e-12
    torch.cuda.synchronize()
    assert rel_err(res.state.position, s.position) <= 1e-12
    assert rel_err(res.state.logdensity, s.logdensity) <= 1e-12
    assert rel_err(res.state.logdensity_grad, s.logdensity_grad) <= 1e-12
    assert int(res.num_fevals.min()) >= T * L * 2


def test_batch_mixed_widths_matches_single_posteriors():
    """Posteriors of 8..128 features in one call: the narrow (<=64) and wide
    (65..128) tilings each take their own chains; every chain must equal the
    same chain run on its posterior alone."""
    shapes = [(200, 8), (700, 24), (333, 100), (1500, 40), (600, 128)]
    Cp, eps, L, T = 2, 0.05, 4, 2
    integ = functools.partial(implicit_midpoint, max_iters=6)
    models = [gj.LogisticRegression(*make_logistic(n, d, 40 + i)[:2], 100.0) for i, (n, d) in enumerate(shapes)]
    batch = gj.LogisticBatch(models, chains_per_posterior=Cp)
    Dmax = max(d for _, d in shapes)
    C = Cp * len(shapes)
    q0 = torch.zeros(C, Dmax, dtype=torch.float64)
    for i, (_, d) in enumerate(shapes):
        q0[i * Cp:(i + 1) * Cp, :d] = 0.1 * seeded_normal(80 + i, Cp, d)
    zs = seeded_normal(90, T, C, Dmax)
    us = seeded_uniform(91, T, C)
    for i, (_, d) in enumerate(shapes):
        zs[:, i * Cp:(i + 1) * Cp, d:] = 0.0
    alg = gj.rmhmc(batch, eps, gj.FisherMetric(), L, integrator=integ)
    s = alg.init(q0.to(DEV))
    for t in range(T):
        s, _ = alg.step(gr.HostNoise(zs[t].to(DEV), us[t].to(DEV)), s)
    torch.cuda.synchronize()
    for i, (m, (_, d)) in enumerate(zip(models, shapes)):
        rows = slice(i * Cp, (i + 1) * Cp)
        print(f"posterior {i}: D={d}", flush=True)
        a1 = gj.rmhmc(m, eps, gj.FisherMetric(), L, integrator=integ)
        s1 = a1.init(q0[rows, :d].contiguous().to(DEV))
        for t in range(T):
            s1, _ = a1.step(gr.HostNoise(zs[t, rows, :d].contiguous().to(DEV), us[t, rows].to(DEV)), s1)
        torch.cuda.synchronize()
        assert rel_err(s.position[rows, :d], s1.position) <= 1e-12
        assert rel_err(s.logdensity[rows], s1.logdensity) <= 1e-12
        if d < Dmax:
            assert float(s.position[rows, d:].abs().max()) == 0.0
