"""GPU: the documented user-facing calls (the Python examples of the integration
guide), each at a small size, so the documented API keeps running as written.
Checks are shapes, dtypes and finiteness; parity lives in the other GPU tests."""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import implicit_midpoint
from geomjax_amd.synthetic import make_logistic

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _finite(*ts):
    for t in ts:
        assert torch.isfinite(t).all()


@pytest.fixture(scope="module")
def logistic():
    X, y, _ = make_logistic(256, 8, 3)
    return gj.LogisticRegression(X, y, prior_var=100.0), 8


def test_rmhmc_funnel_step_run_chains_ess():
    print("case rmhmc funnel float32 + run_chains + ess", flush=True)
    integ = functools.partial(implicit_midpoint, max_iters=8)
    sampler = gj.rmhmc(gj.NealFunnel(D=100), 0.03, gj.MongeMetric(alpha2=1.0), 8, integrator=integ)
    state = sampler.init(torch.randn(256, 100, device=DEV))
    state, info = sampler.step(gr.key(0), state)
    res = gj.sampling.run_chains(sampler, state, 20, gr.key(1))
    ess = gj.ess(res.samples, chain_axis=1, sample_axis=0)
    torch.cuda.synchronize()
    assert res.samples.shape == (20, 256, 100) and ess.shape == (100,)
    _finite(state.position, res.samples)


def test_rmhmc_logistic_sample_fn_and_batch(logistic):
    print("case rmhmc logistic + sample_fn + LogisticBatch", flush=True)
    model, D = logistic
    integ = functools.partial(implicit_midpoint, max_iters=6)
    sampler = gj.rmhmc(model, 0.1, gj.FisherMetric(), 10, integrator=integ)
    state = sampler.init(torch.randn(64, D, device=DEV))
    res = sampler.sample_fn(gr.key(2), state, 5)
    X2, y2, _ = make_logistic(128, 5, 4)
    batch = gj.LogisticBatch([model, gj.LogisticRegression(X2, y2, prior_var=100.0)], 16)
    bs = gj.rmhmc(batch, 0.1, gj.FisherMetric(), 10, integrator=integ)
    bstate, binfo = bs.step(gr.key(3), bs.init(torch.zeros(32, D, device=DEV, dtype=torch.float64)))
    torch.cuda.synchronize()
    _finite(res.samples, bstate.position)


def test_dynamic_rmhmc_and_chees(logistic):
    print("case dynamic_rmhmc + chees_adaptation", flush=True)
    from geomjax_amd.adaptation import adam, chees_adaptation

    model, D = logistic
    C = 64
    positions = 0.1 * torch.randn(C, D, device=DEV, dtype=torch.float64)
    dyn = gj.dynamic_rmhmc(model, 0.1, gj.FisherMetric(), integrator=functools.partial(implicit_midpoint, max_iters=6))
    dstate = dyn.init(positions, torch.arange(C, device=DEV))
    dstate, info = dyn.step(gr.key(5), dstate)
    assert info.num_integration_steps.shape == (C,)
    (last, params), hist = chees_adaptation(model, gj.FisherMetric(), C).run(gr.key(6), positions, 0.1, adam(0.05), 5)
    tuned = gj.dynamic_rmhmc(model, **params)
    ts, _ = tuned.step(gr.key(7), last)  # the adaptation's last state is a DynamicRMHMCState
    torch.cuda.synchronize()
    _finite(dstate.position, ts.position)


def test_nuts_family(logistic):
    print("case nutsrmhmc / nutslmcmonge / nutslmc", flush=True)
    model, D = logistic
    nuts = gj.nutsrmhmc(gj.NealFunnel(D=100), 0.05, gj.MongeMetric(1.0), max_num_doublings=6,
                        integrator=functools.partial(implicit_midpoint, max_iters=8))
    nstate = nuts.init(torch.randn(128, 100, device=DEV, dtype=torch.float64))
    nstate, ninfo = nuts.step(gr.key(7), nstate)
    lnuts = gj.nutsrmhmc(model, 0.1, gj.FisherMetric(), max_num_doublings=5)
    ls, _ = lnuts.step(gr.key(8), lnuts.init(0.1 * torch.randn(16, D, device=DEV, dtype=torch.float64)))
    im = torch.ones(100, dtype=torch.float64, device=DEV)
    mnuts = gj.nutslmcmonge(gj.NealFunnel(D=100), 0.03, im, alpha2=1.0, max_num_doublings=6)
    ms, minfo = mnuts.step(gr.key(9), mnuts.init(torch.randn(128, 100, device=DEV, dtype=torch.float64)))
    fnuts = gj.nutslmc(model, 0.1, gj.FisherMetric(), max_num_doublings=5)
    fs, finfo = fnuts.step(gr.key(10), fnuts.init(0.1 * torch.randn(16, D, device=DEV, dtype=torch.float64)))
    torch.cuda.synchronize()
    assert ninfo.num_integration_steps.shape == (128,) and finfo.velocity.shape == (16, D)
    _finite(nstate.position, ls.position, ms.position, fs.position, fs.volume_adjustment)


def test_lmc_family_and_lmcmonge_chees(logistic):
    print("case lmcmonge / lmc / dynamic_lmc / dynamic_lmcmonge / chees_adaptation_lmcmonge", flush=True)
    from geomjax_amd.adaptation import adam, chees_adaptation_lmcmonge

    model, D = logistic
    im = torch.ones(100, dtype=torch.float64, device=DEV)
    lmc_m = gj.lmcmonge(gj.NealFunnel(D=100), 0.03, im, 8, alpha2=1.0)
    sm, _ = lmc_m.step(gr.key(11), lmc_m.init(torch.randn(128, 100, device=DEV, dtype=torch.float64)))
    lmc_f = gj.lmc(model, 0.1, gj.FisherMetric(), 4)
    sf, _ = lmc_f.step(gr.key(12), lmc_f.init(0.1 * torch.randn(16, D, device=DEV, dtype=torch.float64)))
    dl = gj.dynamic_lmc(model, 0.1, gj.FisherMetric())
    sdl, _ = dl.step(gr.key(13), dl.init(0.1 * torch.randn(16, D, device=DEV, dtype=torch.float64),
                                          torch.arange(16, device=DEV)))
    im10 = torch.ones(10, dtype=torch.float64, device=DEV)
    dlm = gj.dynamic_lmcmonge(gj.NealFunnel(D=10), 0.05, im10, alpha2=0.5)
    sdm, _ = dlm.step(gr.key(14), dlm.init(torch.randn(64, 10, device=DEV, dtype=torch.float64),
                                           torch.arange(64, device=DEV)))
    C = 64
    warm = chees_adaptation_lmcmonge(gj.NealFunnel(D=10), im10, C, update_alpha=True, alpha2=0.5)
    (last, params), hist = warm.run(gr.key(8), torch.randn(C, 10, device=DEV, dtype=torch.float64), 0.05,
                                    adam(0.05), 5)
    tuned = gj.dynamic_lmcmonge(gj.NealFunnel(D=10), **params)
    torch.cuda.synchronize()
    _finite(sm.position, sf.position, sdl.position, sdm.position, last.position)
    assert tuned is not None


def test_mmala(logistic):
    print("case mmala step + sample_fn", flush=True)
    model, D = logistic
    mm = gj.mmala(model, 0.3, gj.FisherMetric(), reference_faithful=False)
    mstate = mm.init(torch.zeros(64, D, dtype=torch.float64, device=DEV))
    mstate, minfo = mm.step(gr.key(9), mstate)
    mres = mm.sample_fn(gr.key(10), mstate, 5)
    torch.cuda.synchronize()
    assert mstate.metric.shape == (64, D, D) and minfo.proposal.Gamma.shape == (64, D)
    _finite(mstate.position, mres.samples, mres.state.metric, mres.state.Gamma)


def test_tangents_and_adapt(logistic):
    print("case esjd_acceptance + adapt", flush=True)
    from geomjax_amd.adjoint import adapt, esjd_acceptance

    model, D = logistic
    s64 = gj.rmhmc.init(0.1 * torch.randn(64, D, device=DEV, dtype=torch.float64), model)
    r = esjd_acceptance(gr.key(3), s64, model, 0.1, gj.FisherMetric(), 6)
    eps, alpha2, s64, hist = adapt(gr.key(4), s64, model, 6, step_size=0.05, num_steps=3)
    torch.cuda.synchronize()
    assert r.d_objective.shape[0] == 64 and eps > 0
    _finite(r.d_objective, s64.position)


def test_graph_sampler(logistic):
    print("case GraphSampler", flush=True)
    from geomjax_amd.sampling import GraphSampler

    model, D = logistic
    sampler = gj.rmhmc(model, 0.1, gj.FisherMetric(), 10, integrator=functools.partial(implicit_midpoint, max_iters=6))
    state = sampler.init(torch.randn(64, D, device=DEV))
    gs = GraphSampler(sampler, state, 4)
    for _ in range(3):
        gs.replay()
    torch.cuda.synchronize()
    _finite(gs.state.position, gs.samples)
