"""GPU: the tuned (float32 state, split-bf16 MFMA) samplers draw from the same
distribution as the float64 ones.

Two checks, both with the device PRNG and independent keys per form (the
forms are not coupled, so agreement is a statement about the laws, not about
shared noise):

* logistic regression (Fisher metric): after T transitions from the same
  initial positions, the tuned and the plain float64 chains are two i.i.d.
  samples (one final state per chain) of the same law; per coordinate the
  means and variances must agree within 5 standard errors (the variance's
  standard error from the sample fourth moment).  Run on the 2048 x 64
  configuration at 4096 chains and on two posteriors of the uneven batch
  (widths 128 and 8) in one batched launch.
* Neal's funnel: chains started from exact draws (v ~ N(0, 9),
  x_i | v ~ N(0, e^v)) must stay exact after ten transitions of the float32
  RMHMC-Monge and LMC-Monge (exact algebra) kernels: v's mean and variance and
  those of the standardised conditionals x_i e^{-v/2} ~ N(0, 1) within 5
  standard errors.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import implicit_midpoint
from geomjax_amd.synthetic import make_logistic, make_uneven_batch
from tests.helpers import seeded_normal

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
K_SE = 5.0


def _moment_z(a, b):
    """Per-coordinate |difference| / standard error of the means and of the
    variances of two independent samples a [n, D], b [m, D]."""
    a, b = a.double().cpu(), b.double().cpu()
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0), b.var(0)
    zm = (ma - mb).abs() / (va / a.shape[0] + vb / b.shape[0]).sqrt()
    m4a, m4b = ((a - ma) ** 4).mean(0), ((b - mb) ** 4).mean(0)
    se_v = ((m4a - va ** 2) / a.shape[0] + (m4b - vb ** 2) / b.shape[0]).clamp_min(1e-300).sqrt()
    zv = (va - vb).abs() / se_v
    return zm, zv


def _final_states(alg, q0, key, T):
    s = alg.init(q0)
    res = alg.sample_fn(key, s, T, store_samples=False)
    torch.cuda.synchronize()
    return res.state.position, res.acceptance_rate.double().mean().item()


def test_tuned_and_plain_moments_agree_logistic_2048x64():
    N, D, C, eps, L, T = 2048, 64, 4096, 0.1, 10, 15
    print(f"case logistic {N}x{D}, {C} chains, {T} transitions", flush=True)
    X, y, _ = make_logistic(N, D, 2)
    model = gj.LogisticRegression(X, y, 100.0)
    alg = gj.rmhmc(model, eps, gj.FisherMetric(), L, integrator=functools.partial(implicit_midpoint, max_iters=6))
    q0 = seeded_normal(301, C, D).to(DEV)
    q32, acc32 = _final_states(alg, q0.float(), gr.key(11), T)
    q64, acc64 = _final_states(alg, q0, gr.key(12), T)
    zm, zv = _moment_z(q32, q64)
    print(f"  acceptance tuned {acc32:.4f} plain {acc64:.4f}; max z of means {zm.max():.2f}, "
          f"of variances {zv.max():.2f}", flush=True)
    assert torch.isfinite(q32).all() and torch.isfinite(q64).all()
    assert abs(acc32 - acc64) < 0.02
    assert float(zm.max()) < K_SE
    assert float(zv.max()) < K_SE


def test_tuned_and_plain_moments_agree_uneven_batch_widths_128_and_8():
    data = make_uneven_batch(P=256, seed=3)
    picks = [0, 1]  # posterior 0: 494 x 128, posterior 1: 1286 x 8
    shapes = [data[i][0].shape for i in picks]
    assert shapes[0][1] == 128 and shapes[1][1] == 8
    cpp, eps, L, T = 2048, 0.1, 10, 15
    models = [gj.LogisticRegression(data[i][0], data[i][1], 100.0) for i in picks]
    batch = gj.LogisticBatch(models, cpp)
    C, Dm = batch.num_chains, batch.dim
    print(f"case uneven-batch posteriors {shapes}, {cpp} chains each, {T} transitions", flush=True)
    q0 = torch.zeros(C, Dm, dtype=torch.float64)
    for p, (_, d) in enumerate(shapes):
        q0[p * cpp:(p + 1) * cpp, :d] = seeded_normal(310 + p, cpp, d)
    alg = gj.rmhmc(batch, eps, gj.FisherMetric(), L, integrator=functools.partial(implicit_midpoint, max_iters=6))
    q32, acc32 = _final_states(alg, q0.float().to(DEV), gr.key(21), T)
    q64, acc64 = _final_states(alg, q0.to(DEV), gr.key(22), T)
    assert torch.isfinite(q32).all() and torch.isfinite(q64).all()
    for p, (_, d) in enumerate(shapes):
        rows = slice(p * cpp, (p + 1) * cpp)
        zm, zv = _moment_z(q32[rows, :d], q64[rows, :d])
        print(f"  posterior {picks[p]} ({shapes[p][0]}x{d}): max z of means {zm.max():.2f}, "
              f"of variances {zv.max():.2f}", flush=True)
        assert float(zm.max()) < K_SE
        assert float(zv.max()) < K_SE
        assert torch.all(q32[rows, d:] == 0)
    print(f"  acceptance tuned {acc32:.4f} plain {acc64:.4f}", flush=True)
    assert abs(acc32 - acc64) < 0.02


def _funnel_exact_draws(D, C, seed):
    g = torch.Generator().manual_seed(seed)
    v = 3.0 * torch.randn(C, 1, generator=g, dtype=torch.float64)
    x = torch.randn(C, D - 1, generator=g, dtype=torch.float64) * torch.exp(0.5 * v)
    return torch.cat([x, v], dim=1)


def _check_funnel_stationary(q, label):
    q = q.double().cpu()
    C = q.shape[0]
    v = q[:, -1]
    std = q[:, :4] * torch.exp(-0.5 * v)[:, None]  # x_i e^{-v/2} ~ N(0, 1), independent of v
    rows = [("v", v, 0.0, 9.0)] + [(f"x{i}*exp(-v/2)", std[:, i], 0.0, 1.0) for i in range(4)]
    for name, x, m0, v0 in rows:
        m, var = float(x.mean()), float(x.var())
        zm = abs(m - m0) / (v0 / C) ** 0.5
        zv = abs(var - v0) / (v0 * (2.0 / C) ** 0.5)
        print(f"  {label} {name}: mean {m:.4f} ({m0}), var {var:.4f} ({v0}); z {zm:.2f}, {zv:.2f}", flush=True)
        assert zm < K_SE and zv < K_SE


def test_rmhmc_monge_f32_keeps_the_funnel_stationary():
    D, C, T = 10, 8192, 10
    print(f"case funnel D={D}, {C} chains, RMHMC-Monge float32, {T} transitions", flush=True)
    model = gj.NealFunnel(D)
    alg = gj.rmhmc(model, 0.05, gj.MongeMetric(1.0), 16, integrator=functools.partial(implicit_midpoint, max_iters=8))
    q, acc = _final_states(alg, _funnel_exact_draws(D, C, 81).float().to(DEV), gr.key(31), T)
    print(f"  acceptance {acc:.3f}", flush=True)
    assert q.dtype == torch.float32 and torch.isfinite(q).all()
    _check_funnel_stationary(q, "rmhmc f32")


def test_lmcmonge_f32_exact_algebra_keeps_the_funnel_stationary():
    D, C, T = 10, 8192, 10
    print(f"case funnel D={D}, {C} chains, LMC-Monge float32 (exact algebra), {T} transitions", flush=True)
    model = gj.NealFunnel(D)
    im = torch.ones(D, dtype=torch.float32, device=DEV)
    alg = gj.lmcmonge(model, 0.05, im, 16, alpha2=1.0, reference_faithful=False)
    q, acc = _final_states(alg, _funnel_exact_draws(D, C, 82).float().to(DEV), gr.key(32), T)
    print(f"  acceptance {acc:.3f}", flush=True)
    assert q.dtype == torch.float32 and torch.isfinite(q).all()
    _check_funnel_stationary(q, "lmcmonge f32")
