"""GPU: the dynamic RMHMC kernel (per-chain numbers of integration steps,
geomjax/rmhmc/rmhmc.py:179-244).

* float64, per-chain counts: every chain equals the oracle run on that chain
  alone with its own count (1e-10 norm-wise relative, identical fixed-point
  iteration counts over the chain's own steps);
* float32 kernels (funnel, one-wave and four-wave logistic) against the
  float64 kernel with the same counts (1e-4 funnel, 1e-5 logistic);
* a persistent launch with [T, C] counts equals T step launches;
* ``dynamic_rmhmc`` draws its counts from the chains' arguments and equals the
  base kernel given those counts.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import _integration_steps, build_kernel, implicit_midpoint
from geomjax_amd.synthetic import make_logistic
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _problem(kind):
    if kind == "funnel":
        D = 10
        M = om.NealFunnel(D)
        return (gj.NealFunnel(D), gj.MongeMetric(1.0), M, om.monge_metric_fn(M.logp, 1.0, D), D, 0.05, 8)
    N, D = (300, 13) if kind == "logistic" else (200, 72)
    X, y, _ = make_logistic(N, D, 31)
    M = om.LogisticRegression(X, y, 100.0)
    return gj.LogisticRegression(X, y, 100.0), gj.FisherMetric(), M, M.fisher_metric_fn, D, 0.05, 6


@pytest.mark.parametrize("kind", ["funnel", "logistic", "logistic_wide"])
def test_per_chain_steps_f64_matches_oracle_chain_by_chain(kind):
    model, metric, M, ometric, D, eps, maxit = _problem(kind)
    steps = torch.tensor([0, 1, 3, 5, 8, 2], dtype=torch.int32)
    C = steps.numel()
    print(f"case {kind}: D={D} steps={steps.tolist()}", flush=True)
    q0 = 0.5 * seeded_normal(41, C, D)
    z, u = seeded_normal(42, C, D), seeded_uniform(43, C)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=maxit), return_fixed_point_info=True)
    s = gj.rmhmc.init(q0.to(DEV), model)
    s1, info, fp = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, model, eps, metric, steps)
    torch.cuda.synchronize()
    assert torch.equal(info.num_integration_steps.cpu(), steps)
    okern = orm.build_kernel(max_iters=maxit)
    for c in range(C):
        Lc = int(steps[c])
        tr = []
        so, io = okern((z[c:c + 1], u[c:c + 1]), orm.init(q0[c:c + 1], M.logp), M.logp, eps, ometric, Lc, trace=tr)
        assert rel_err(s1.position[c:c + 1], so.position) < 1e-10, c
        assert rel_err(info.proposal.state.position[c:c + 1], io.proposal.state.position) < 1e-10, c
        assert abs(float(info.energy[c]) - float(io.energy[0])) <= 1e-10 * max(1.0, abs(float(io.energy[0]))), c
        assert bool(info.is_accepted[c]) == bool(io.is_accepted[0]), c
        if Lc:
            assert torch.equal(fp.iters[c, :Lc].cpu().long(), torch.stack([x.iters for x in tr], dim=1)[0]), c


@pytest.mark.parametrize("kind", ["funnel", "logistic", "logistic_wide"])
def test_per_chain_steps_f32_matches_f64(kind):
    model, metric, _, _, D, eps, maxit = _problem(kind)
    C = 12
    steps = torch.tensor([(3 * c) % 9 for c in range(C)], dtype=torch.int32)
    q0 = 0.5 * seeded_normal(51, C, D)
    z, u = seeded_normal(52, C, D), seeded_uniform(53, C)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=maxit))
    s64, i64 = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), gj.rmhmc.init(q0.to(DEV), model), model, eps, metric, steps)
    s32, i32 = kern(gr.HostNoise(z.to(DEV).float(), u.to(DEV)), gj.rmhmc.init(q0.to(DEV).float(), model), model,
                    eps, metric, steps)
    torch.cuda.synchronize()
    tol = 1e-4 if kind == "funnel" else 1e-5
    e = rel_err(i32.proposal.state.position, i64.proposal.state.position)
    print(f"case {kind}: proposal position f32 vs f64 {e:.2e}", flush=True)
    assert e < tol
    assert (i32.proposal.weight.double() - i64.proposal.weight).abs().max().item() < 1e-3


@pytest.mark.parametrize("kind,dtype", [("funnel", torch.float64), ("logistic", torch.float64),
                                        ("funnel", torch.float32), ("logistic", torch.float32),
                                        ("logistic_wide", torch.float32)])
def test_sample_with_per_transition_steps_matches_steps(kind, dtype):
    from geomjax_amd.rmhmc import sample

    model, metric, _, _, D, eps, maxit = _problem(kind)
    C, T = 8, 4
    steps = torch.tensor([[(c + 2 * t) % 7 for c in range(C)] for t in range(T)], dtype=torch.int32)
    integ = functools.partial(implicit_midpoint, max_iters=maxit)
    q0 = (0.5 * seeded_normal(61, C, D)).to(DEV, dtype)
    zs, us = seeded_normal(62, T, C, D).to(DEV, dtype), seeded_uniform(63, T, C).to(DEV)
    s = gj.rmhmc.init(q0, model)
    r = sample(gr.HostNoise(zs, us), s, model, eps, metric, steps, T, integrator=integ)
    kern = build_kernel(integ)
    for t in range(T):
        s, _ = kern(gr.HostNoise(zs[t], us[t]), s, model, eps, metric, steps[t])
        assert rel_err(r.samples[t], s.position) <= (1e-12 if dtype == torch.float64 else 1e-6), t
    torch.cuda.synchronize()


def test_dynamic_rmhmc_api():
    model, metric, _, _, D, eps, maxit = _problem("logistic")
    C = 16
    integ = functools.partial(implicit_midpoint, max_iters=maxit)
    alg = gj.dynamic_rmhmc(model, eps, metric, integrator=integ)
    state = alg.init((0.3 * seeded_normal(71, C, D)).to(DEV), torch.arange(C, device=DEV) * 7919)
    base = build_kernel(integ)
    for t in range(3):
        z, u = seeded_normal(72 + t, C, D).to(DEV), seeded_uniform(82 + t, C).to(DEV)
        expect = _integration_steps(state.random_generator_arg)
        assert int(expect.min()) >= 1 and int(expect.max()) <= 9
        new, info = alg.step(gr.HostNoise(z, u), state)
        ref, _ = base(gr.HostNoise(z, u), gj.rmhmc.init(state.position, model), model, eps, metric, expect)
        torch.cuda.synchronize()
        assert torch.equal(info.num_integration_steps, expect)
        assert rel_err(new.position, ref.position) <= 1e-12
        assert torch.equal(new.random_generator_arg, state.random_generator_arg + 1)
        state = new
    # the persistent launch draws the same counts as the step loop
    zs, us = seeded_normal(90, 3, C, D).to(DEV), seeded_uniform(91, 3, C).to(DEV)
    r = alg.sample_fn(gr.HostNoise(zs, us), state, 3)
    s = state
    for t in range(3):
        s, _ = alg.step(gr.HostNoise(zs[t], us[t]), s)
        assert rel_err(r.samples[t], s.position) <= 1e-12
    assert torch.equal(r.state.random_generator_arg, s.random_generator_arg)
