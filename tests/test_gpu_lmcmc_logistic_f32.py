"""GPU parity: tuned LMC with the Fisher metric (float32 state, split-bf16 MFMA
Gram matrices, float64 log-determinants, volume change and energies) on
Bayesian logistic regression.

* one transition against the float64 oracle (jacfwd Omega~, LU determinants)
  from the same float32-representable start and the same draws: velocity,
  proposal position, momentum and gradient within 1e-4 norm-wise relative
  (observed <= 3e-5), the volume adjustment (2L float64 sums of float32
  log-determinant differences) within 2e-4 absolute (observed <= 2e-5),
  energy and dH within 1e-2 absolute (observed <= 5e-3 at 2048 x 64: the
  float32 velocity's ~2e-5 relative error enters v^T G v ~ D), accept
  decisions equal outside |log u - dH| < 1e-2;
* the persistent launch equals step launches (1e-5: the two forms are
  compiled separately);
* a batch of posteriors of both tile widths equals each posterior alone;
* paired with the float64 kernel on the device PRNG over 30 transitions, the
  chains' means agree within 5 standard errors and the acceptance rates within
  0.02.
"""
import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.lmcmc import build_kernel
from geomjax_amd.synthetic import make_logistic
from oracle import lmcmc as olc
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("N,D,C,eps,L,seed", [
    (50, 6, 3, 0.1, 5, 5),       # one 64-row group: the smallest row loops
    (256, 8, 3, 0.1, 6, 0),
    (300, 13, 2, 0.1, 4, 1),
    (1024, 32, 2, 0.1, 5, 2),    # the one-block width's capacity
    (1024, 40, 2, 0.1, 5, 3),    # the two-block width
    (2048, 64, 2, 0.1, 6, 4),    # the kernel's feature capacity, on the 2048 x 64 shape
])
def test_tuned_step_matches_oracle(N, D, C, eps, L, seed):
    print(f"case tuned LMC N={N} D={D} C={C} eps={eps} L={L}", flush=True)
    X, y, _ = make_logistic(N, D, 10 + seed)
    M = om.LogisticRegression(X, y, 100.0)
    q0 = (seeded_normal(seed, C, D) * 0.5).float().double()
    z = seeded_normal(700 + seed, C, D).float().double()
    u = seeded_uniform(800 + seed, C)
    so = olc.init(q0, M.logp)
    so1, io = olc.build_kernel()((z, u), so, M.logp, eps, M.fisher_metric_fn, L)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.lmc.init(q0.float().to(DEV), model)
    s1, info = build_kernel()(gr.HostNoise(z.float().to(DEV), u.to(DEV)), s, model, eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    assert s1.position.dtype == torch.float32
    ps, pso = info.proposal.state, io.proposal.state
    errs = {name: rel_err(getattr(ps, name).double(), getattr(pso, name))
            for name in ("position", "momentum", "velocity", "logdensity_grad")}
    errs["v0"] = rel_err(info.velocity.double(), io.velocity)
    de = (info.energy.cpu() - io.energy).abs().max().item()
    dj = (ps.volume_adjustment.cpu() - pso.volume_adjustment).abs().max().item()
    dw = (info.proposal.weight.cpu() - io.proposal.weight).abs()
    print(f"  errors {', '.join(f'{k} {v:.1e}' for k, v in errs.items())}; energy {de:.1e}, J {dj:.1e}, "
          f"weight {dw.max().item():.1e}; J {pso.volume_adjustment.tolist()}", flush=True)
    for name, e in errs.items():
        assert e < 1e-4, (name, e)
    assert dj < 2e-4
    assert de < 1e-2 and dw.max().item() < 1e-2
    ties = (torch.log(u) - io.proposal.weight).abs() < 1e-2
    assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])
    assert rel_err(s1.position.cpu()[~ties].double(), so1.position[~ties]) < 1e-4


def test_tuned_sample_equals_step_launches():
    from geomjax_amd.sampling import run_chains

    N, D, C, eps, L, T = 512, 40, 24, 0.1, 4, 3
    X, y, _ = make_logistic(N, D, 17)
    alg = gj.lmc(gj.LogisticRegression(X, y, 100.0), eps, gj.FisherMetric(), L)
    s = alg.init((seeded_normal(81, C, D) * 0.5).float().to(DEV))
    zs, us = seeded_normal(82, T, C, D).float().to(DEV), seeded_uniform(83, T, C).to(DEV)
    res = run_chains(alg, s, T, gr.HostNoise(zs, us))
    for t in range(T):
        s, info = alg.step(gr.HostNoise(zs[t], us[t]), s)
        torch.cuda.synchronize()
        err = rel_err(res.samples[t].double(), s.position.double())
        print(f"transition {t}: accepted {int(info.is_accepted.sum())}/{C}, rel err {err:.1e}", flush=True)
        assert err < 1e-5
    assert res.samples.dtype == torch.float32
    assert rel_err(res.state.position.double(), s.position.double()) < 1e-5
    assert rel_err(res.state.logdensity_grad.double(), s.logdensity_grad.double()) < 1e-4


def test_tuned_batch_equals_each_posterior_alone():
    shapes = [(300, 8), (500, 40), (256, 24)]
    models = [gj.LogisticRegression(*make_logistic(n, d, 40 + i)[:2], 100.0) for i, (n, d) in enumerate(shapes)]
    cpp = 4
    batch = gj.LogisticBatch(models, chains_per_posterior=cpp)
    Dm = batch.dim
    C = cpp * len(models)
    q = torch.zeros(C, Dm)
    z = torch.zeros(C, Dm)
    for i, (_, d) in enumerate(shapes):
        q[i * cpp:(i + 1) * cpp, :d] = 0.3 * seeded_normal(50 + i, cpp, d)
        z[i * cpp:(i + 1) * cpp, :d] = seeded_normal(60 + i, cpp, d)
    u = seeded_uniform(70, C)
    kern = build_kernel()
    sb = gj.lmc.init(q.float().to(DEV), batch)
    sb1, ib = kern(gr.HostNoise(z.float().to(DEV), u.to(DEV)), sb, batch, 0.1, gj.FisherMetric(), 4)
    for i, (_, d) in enumerate(shapes):
        rows = slice(i * cpp, (i + 1) * cpp)
        s = gj.lmc.init(q[rows, :d].float().to(DEV), models[i])
        s1, info = kern(gr.HostNoise(z[rows, :d].float().to(DEV), u[rows].to(DEV)), s, models[i], 0.1,
                        gj.FisherMetric(), 4)
        torch.cuda.synchronize()
        err = rel_err(sb1.position[rows, :d].double(), s1.position.double())
        print(f"posterior {i} (D={d}): rel err {err:.1e}", flush=True)
        assert err < 1e-6
        assert torch.equal(ib.is_accepted[rows], info.is_accepted)
        assert (ib.energy[rows] - info.energy).abs().max().item() < 1e-6 * max(info.energy.abs().max().item(), 1.0)
        assert torch.all(sb1.position[rows, d:] == 0)


def test_tuned_chains_track_float64_chains():
    N, D, C, eps, L, T = 512, 8, 2048, 0.15, 6, 30
    X, y, _ = make_logistic(N, D, 23)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = 0.1 * seeded_normal(91, C, D)
    alg = gj.lmc(model, eps, gj.FisherMetric(), L)
    out = {}
    for dt in (torch.float64, torch.float32):
        res = alg.sample_fn(gr.PRNGKey(92, 0), alg.init(q0.to(DEV).to(dt)), T)
        torch.cuda.synchronize()
        out[dt] = (res.state.position.double().cpu(), float(res.acceptance_rate.double().mean()))
    (p64, a64), (p32, a32) = out[torch.float64], out[torch.float32]
    se = p64.std(0) / C ** 0.5
    dm = (p32.mean(0) - p64.mean(0)).abs()
    print(f"acceptance {a64:.4f} (float64) vs {a32:.4f} (tuned); mean diff / SE max {(dm / se).max().item():.2f}; "
          f"chains apart (rel) {rel_err(p32, p64):.1e}", flush=True)
    assert abs(a64 - a32) < 0.02
    assert torch.all(dm < 5 * se)
