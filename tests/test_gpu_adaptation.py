"""GPU: the ChEES driver (geomjax_amd.adaptation.chees_adaptation) over the HIP
RMHMC kernel: the step size moves the harmonic-mean acceptance towards the
0.651 target, every parameter stays finite, and the adapted parameters build
a working dynamic RMHMC sampler; the same for the LMC-Monge driver
(chees_adaptation_lmcmonge) with its alpha2 update, whose per-iteration
derivative comes from the alpha2 tangent kernel.  (The drivers' arithmetic is pinned on the
CPU in tests/test_adaptation.py; sampling noise makes the trajectory of the
adaptation itself a statistical, not an exact, check.)"""
import functools
import math

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.adaptation import adam, chees_adaptation
from geomjax_amd.rmhmc import implicit_midpoint
from geomjax_amd.synthetic import make_logistic
from tests.helpers import seeded_normal

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("kind", ["funnel", "logistic"])
def test_chees_adapts_and_the_result_samples(kind):
    if kind == "funnel":
        D, model, metric, eps0 = 10, gj.NealFunnel(10), gj.MongeMetric(1.0), 0.5
    else:
        X, y, _ = make_logistic(400, 8, 21)
        D, model, metric, eps0 = 8, gj.LogisticRegression(X, y, 100.0), gj.FisherMetric(), 0.1
    C = 128
    integ = functools.partial(implicit_midpoint, max_iters=8)
    warm = chees_adaptation(model, metric, C, integrator=integ)
    q0 = (0.3 * seeded_normal(9, C, D)).to(DEV)
    (last, params), hist = warm.run(gr.key(11), q0, eps0, adam(0.05), num_steps=60)
    torch.cuda.synchronize()
    steps = [h[0] for h in hist]
    accs = [h[3] for h in hist]
    harm = [h[4] for h in hist]  # the statistic the dual averaging drives to 0.651
    print(f"{kind}: step size {steps[0]:.3f} -> {params['step_size']:.3f}, "
          f"L range {min(h[2] for h in hist)}..{max(h[2] for h in hist)}, "
          f"mean acceptance first 10 {sum(accs[:10]) / 10:.3f} last 20 {sum(accs[-20:]) / 20:.3f}, "
          f"harmonic first 10 {sum(harm[:10]) / 10:.3f} last 20 {sum(harm[-20:]) / 20:.3f}", flush=True)
    assert all(math.isfinite(h[0]) and math.isfinite(h[1]) for h in hist)
    assert math.isfinite(params["step_size"]) and params["step_size"] > 0
    assert params["max_num_integration_steps"] >= 1
    # the last 20 iterations sit closer to the target than the first 10
    assert abs(sum(harm[-20:]) / 20 - 0.651) < abs(sum(harm[:10]) / 10 - 0.651) + 0.1
    alg = gj.dynamic_rmhmc(model, **params)
    state = last
    for t in range(5):
        state, info = alg.step(gr.key(100 + t), state)
    torch.cuda.synchronize()
    assert torch.isfinite(state.position).all()
    assert float(info.acceptance_rate.mean()) > 0.2



@pytest.mark.parametrize("faithful", [True, False])
def test_chees_lmcmonge_adapts_and_the_result_samples(faithful):
    from geomjax_amd.adaptation import chees_adaptation_lmcmonge

    D, C = 10, 128
    model = gj.NealFunnel(D)
    im = torch.ones(D, dtype=torch.float64, device=DEV)
    warm = chees_adaptation_lmcmonge(model, im, C, update_alpha=True, alpha2=0.5, reference_faithful=faithful)
    q0 = (0.3 * seeded_normal(19, C, D)).to(DEV)
    (last, params), hist = warm.run(gr.key(21), q0, 0.05, adam(0.05), num_steps=40)
    torch.cuda.synchronize()
    harm = [h[4] for h in hist]
    a2s = [h[5] for h in hist]
    print(f"lmcmonge faithful={faithful}: step size {hist[0][0]:.3f} -> {params['step_size']:.3f}, "
          f"L {min(h[2] for h in hist)}..{max(h[2] for h in hist)}, harmonic first 5 {sum(harm[:5]) / 5:.3f} "
          f"last 15 {sum(harm[-15:]) / 15:.3f}, alpha2 {a2s[0]:.3f} -> {params['alpha2']:.3f}", flush=True)
    assert all(math.isfinite(h[0]) and math.isfinite(h[1]) and math.isfinite(h[5]) for h in hist)
    assert math.isfinite(params["step_size"]) and params["step_size"] > 0
    assert math.isfinite(params["alpha2"]) and params["alpha2"] > 0
    assert a2s[0] == (1.0 if faithful else 0.5)
    assert params["alpha2"] != a2s[0]  # the alpha2 update ran
    assert last.alpha2 == 0.5 or not faithful  # the faithful driver never writes alpha2 back to the chains
    alg = gj.dynamic_lmcmonge(model, **params)
    state = last
    for t in range(5):
        state, info = alg.step(gr.key(200 + t), state)
    torch.cuda.synchronize()
    assert torch.isfinite(state.position).all()
