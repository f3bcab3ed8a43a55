"""GPU: the ChEES driver (geomjax_amd.adaptation.chees_adaptation) over the HIP
RMHMC kernel: the step size moves the harmonic-mean acceptance towards the
0.651 target, every parameter stays finite, and the adapted parameters build
a working dynamic RMHMC sampler.  (The driver's arithmetic is pinned on the
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

