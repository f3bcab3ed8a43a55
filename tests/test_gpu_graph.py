"""GPU: persistent sample launches captured into a HIP graph (sampling.GraphSampler).

Each case captures a chunk of T transitions once and replays it several
times, with the program's own eager work (filling the noise buffers, an eager
run of the same chunk, synchronisation) between the replays; every replay
must equal the eager launch on the same noise (same kernel, same inputs:
1e-12 norm-wise relative).  With the noise drawn inside the graph, successive
replays must advance the chains with fresh draws.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import implicit_midpoint
from geomjax_amd.sampling import GraphSampler
from geomjax_amd.synthetic import make_logistic
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _cases():
    integ = functools.partial(implicit_midpoint, max_iters=6)
    X, y, _ = make_logistic(512, 40, 11)
    Xw, yw, _ = make_logistic(300, 100, 12)
    return {
        "logistic_f32": lambda: (gj.rmhmc(gj.LogisticRegression(X, y, 100.0), 0.1, gj.FisherMetric(), 6,
                                          integrator=integ), 48, 40, torch.float32),
        "logistic_wide_f32": lambda: (gj.rmhmc(gj.LogisticRegression(Xw, yw, 100.0), 0.05, gj.FisherMetric(), 4,
                                               integrator=integ), 16, 100, torch.float32),
        "funnel_f32": lambda: (gj.rmhmc(gj.NealFunnel(20), 0.05, gj.MongeMetric(1.0), 8,
                                        integrator=functools.partial(implicit_midpoint, max_iters=8)),
                               256, 20, torch.float32),
        "lmcmonge_f32": lambda: (gj.lmcmonge(gj.NealFunnel(20), 0.05, torch.ones(20, device=DEV), 8, alpha2=1.0),
                                 256, 20, torch.float32),
        "logistic_f64": lambda: (gj.rmhmc(gj.LogisticRegression(X, y, 100.0), 0.1, gj.FisherMetric(), 6,
                                          integrator=integ), 16, 40, torch.float64),
        # a user dense metric: torch.func derivatives and the dense-metric kernels in one graph
        "dense_lmc_f64": lambda: (gj.lmc(_funnel_logp, 0.05, _monge_metric, 3), 16, 10, torch.float64),
        # a batch of three widths: its launches fork onto side streams and join
        "logistic_batch_f32": lambda: (gj.rmhmc(_batch(), 0.05, gj.FisherMetric(), 4, integrator=integ), 24, 100,
                                       torch.float32),
    }


def _funnel_logp(q):
    x, v = q[:-1], q[-1]
    return -0.5 * v * v / 9.0 + torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v)


def _monge_metric(q):
    g = torch.func.grad(_funnel_logp)(q)
    return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + torch.outer(g, g)


def _batch():
    models = []
    for N, D, seed in ((200, 20, 13), (300, 40, 14), (250, 100, 15)):
        Xb, yb, _ = make_logistic(N, D, seed)
        models.append(gj.LogisticRegression(Xb, yb, 100.0))
    return gj.LogisticBatch(models, 8)


@pytest.mark.parametrize("name", ["logistic_f32", "logistic_wide_f32", "funnel_f32", "lmcmonge_f32", "logistic_f64",
                                  "logistic_batch_f32", "dense_lmc_f64"])
def test_graph_replays_match_eager(name):
    alg, C, D, dt = _cases()[name]()
    T, R = 3, 3
    print(f"case {name}: C={C} D={D} T={T} replays={R}", flush=True)
    state = alg.init((0.3 * seeded_normal(5, C, D)).to(DEV, dt))
    gs = GraphSampler(alg, state, T, draw_noise=False)
    eager = state
    for rep in range(R):
        z = seeded_normal(100 + rep, T, C, D).to(DEV, dt)
        u = seeded_uniform(200 + rep, T, C).to(DEV)
        gs.z.copy_(z)
        gs.u.copy_(u)
        gs.replay()
        r = alg.sample_fn(gr.HostNoise(z, u), eager, T)  # eager work between replays
        eager = r.state
        torch.cuda.synchronize()
        print(f"  replay {rep}: accepted {int(r.num_accepted.sum())}", flush=True)
        assert rel_err(gs.samples, r.samples) <= 1e-12
        for a, b in zip(gs.state, eager):
            assert rel_err(a, b) <= 1e-12
        assert torch.equal(gs.acceptance_rate, r.acceptance_rate)


def test_graph_draws_fresh_noise_each_replay():
    alg, C, D, dt = _cases()["logistic_f32"]()
    state = alg.init((0.3 * seeded_normal(6, C, D)).to(DEV, dt))
    gs = GraphSampler(alg, state, 4)
    zs, qs = [], []
    for _ in range(3):
        gs.replay()
        torch.cuda.synchronize()
        zs.append(gs.z.clone())
        qs.append(gs.state.position.clone())
    assert not torch.equal(zs[0], zs[1]) and not torch.equal(zs[1], zs[2])
    assert not torch.equal(qs[0], qs[1])
    assert all(torch.isfinite(q).all() for q in qs)
    assert abs(float(zs[2].double().mean())) < 0.1 and abs(float(zs[2].double().std()) - 1.0) < 0.1
    assert 0.3 < float(gs.acceptance_rate.mean()) <= 1.0
