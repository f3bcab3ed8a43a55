"""GPU parity: generalized LMC (glmc: persistent velocity, non-reversible
slice) on the dense-metric path against the glmc oracle (oracle/glmc.py, the
reference's algorithm with its IntegratorState construction corrected), on
the same draws.

Cases: Neal's funnel with the Monge metric built by autodiff (alpha 0.3 and
1, slice noise on and off) and logistic regression with the Fisher metric
given as the model descriptor.  Two transitions are chained on the device.
Bars: accept decisions identical; position, velocity, log density, gradient,
slice, volume adjustment, energy and the proposal 1e-10 norm-wise relative.
"""
import math

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.glmc import GLMCState, build_kernel, init
from geomjax_amd.synthetic import make_logistic
from oracle import glmc as og
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def funnel_logp(q, sigma=3.0):
    x, v = q[:-1], q[-1]
    lv = -0.5 * v * v / sigma ** 2 - 0.5 * math.log(2 * math.pi * sigma ** 2)
    lx = torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi))
    return lv + lx


def monge(alpha2):
    g = torch.func.grad(funnel_logp)

    def metric_fn(q):
        gq = g(q)
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + alpha2 * torch.outer(gq, gq)

    return metric_fn


def _compare(st, info, so, oi, label):
    print(f"  {label}: accepted {oi.is_accepted.tolist()}", flush=True)
    assert info.is_accepted.cpu().tolist() == oi.is_accepted.tolist()
    for f in GLMCState._fields:
        e = rel_err(getattr(st, f), getattr(so, f))
        assert e < 1e-10, (label, f, e)
    assert rel_err(info.energy, oi.energy) < 1e-10
    assert rel_err(info.velocity, oi.velocity) < 1e-10
    for f in ("position", "momentum", "velocity", "logdensity_grad"):
        assert rel_err(getattr(info.proposal.state, f), getattr(oi.proposal.state, f)) < 1e-10, f


@pytest.mark.parametrize("D,C,eps,alpha,delta,noisy,seed", [
    (6, 4, 0.1, 0.3, 0.37, False, 1),
    (10, 3, 0.05, 1.0, 0.5, True, 2),
])
def test_funnel_monge_matches_oracle(D, C, eps, alpha, delta, noisy, seed):
    print(f"case glmc funnel-Monge D={D} C={C} alpha={alpha} delta={delta} noise={noisy}", flush=True)
    metric = monge(1.0)
    noise_fn = (lambda u: 0.1 * u) if noisy else None
    q0 = 0.5 * seeded_normal(10 + seed, C, D)
    z0, u0 = seeded_normal(11 + seed, C, D), seeded_uniform(12 + seed, C)
    s = init(q0.to(DEV), gr.HostNoise(z0.to(DEV), u0.to(DEV)), funnel_logp)
    so = og.init(q0, z0, u0, funnel_logp)
    for f in GLMCState._fields:
        assert rel_err(getattr(s, f), getattr(so, f)) < 1e-12, ("init", f)
    kern, okern = build_kernel(noise_fn), og.build_kernel()
    for t in range(2):
        z, u = seeded_normal(20 + 10 * seed + t, C, D), seeded_uniform(30 + 10 * seed + t, C)
        s, info = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, funnel_logp, eps, metric, alpha, delta)
        so, oi = okern((z, 0.1 * u if noisy else torch.zeros(C, dtype=torch.float64)), so, funnel_logp, eps, metric,
                       alpha, delta)
        torch.cuda.synchronize()
        _compare(s, info, so, oi, f"t={t}")


def test_logistic_fisher_descriptor_matches_oracle():
    N, D, C, eps, alpha, delta = 200, 8, 4, 0.3, 0.5, 0.25
    X, y, _ = make_logistic(N, D, 9)
    M = om.LogisticRegression(X, y, 100.0)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = 0.3 * seeded_normal(40, C, D)
    z0, u0 = seeded_normal(41, C, D), seeded_uniform(42, C)
    s = init(q0.to(DEV), gr.HostNoise(z0.to(DEV), u0.to(DEV)), model)
    so = og.init(q0, z0, u0, M.logp)
    kern, okern = build_kernel(), og.build_kernel()
    for t in range(2):
        z, u = seeded_normal(43 + t, C, D), seeded_uniform(45 + t, C)
        s, info = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, model, eps, gj.FisherMetric(), alpha, delta)
        so, oi = okern((z, torch.zeros(C, dtype=torch.float64)), so, M.logp, eps, M.fisher_metric_fn, alpha, delta)
        torch.cuda.synchronize()
        _compare(s, info, so, oi, f"logistic t={t}")


def test_sampling_algorithm_with_device_draws():
    alg = gj.glmc(funnel_logp, 0.05, monge(1.0), 0.5, 0.3)
    s = alg.init(0.5 * seeded_normal(50, 64, 8).to(DEV), gr.key(1))
    res = alg.sample_fn(gr.key(2), s, 6)
    one, _ = alg.step(gr.key(3), s)
    torch.cuda.synchronize()
    assert res.samples.shape == (6, 64, 8) and torch.isfinite(res.samples).all()
    assert torch.isfinite(one.position).all() and (one.slice.abs() <= 1.0).all()
    assert 0.0 < float(res.acceptance_rate.mean()) <= 1.0
