"""GPU: NUTS with the RMHMC integrator (geomjax_amd.nuts) on Neal's funnel.

* parity with the float64 oracle (oracle/nuts.py) fed the same draws: new
  position, log density, energy, acceptance statistic and the trajectory's
  ends within 1e-10 norm-wise relative; numbers of integration steps and of
  expansions, and the turning / divergence flags identical;
* the same for the "bet" and "riem" criteria, for logistic regression with the
  Fisher metric (and a batch of posteriors against each alone), and for the
  LMC-Monge integrator (geomjax_amd.lmcmonge_nuts) with the volume adjustment
  and both algebras;
* stationarity with the device PRNG (RMHMC, and LMC-Monge in the exact algebra): chains started from exact draws of the
  funnel (v ~ N(0, 9), x_i | v ~ N(0, e^v)) keep v's mean and variance after
  ten transitions (within 5 standard errors).
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.nuts import NUTSNoise
from geomjax_amd.nuts import build_kernel as nuts_kernel
from geomjax_amd.rmhmc import RMHMCState, implicit_midpoint
from oracle import models as om
from oracle import nuts as onuts
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("D,C,eps,M,T,seed,crit", [
    (10, 4, 0.1, 5, 3, 0, "euc"),     # the 10-dimensional funnel, three chained transitions
    (2, 6, 0.3, 6, 2, 1, "euc"),      # minimal funnel, long trajectories
    (100, 2, 0.05, 4, 1, 2, "euc"),   # two coordinates per lane
    (4, 4, 0.15, 5, 2, 3, "bet"),     # the other U-turn criteria
    (4, 4, 0.15, 5, 2, 3, "riem"),
    (160, 2, 0.04, 4, 1, 4, "riem"),  # three (of four) coordinates per lane
])
def test_nuts_matches_oracle(D, C, eps, M, T, seed, crit):
    print(f"case D={D} C={C} eps={eps} M={M} T={T} criterion={crit}", flush=True)
    model = gj.NealFunnel(D)
    Mo = om.NealFunnel(D)
    metric = om.monge_metric_fn(Mo.logp, 1.0, D)
    q0 = seeded_normal(10 + seed, C, D)
    so = orm.init(q0, Mo.logp)
    s = RMHMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV))
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=8))
    okern = onuts.build_kernel(max_iters=8)
    for t in range(T):
        noise = (seeded_normal(20 + 10 * seed + t, C, D), seeded_uniform(30 + 10 * seed + t, C, M),
                 seeded_uniform(40 + 10 * seed + t, C, M), seeded_uniform(50 + 10 * seed + t, C, 2 ** M - 1))
        s1, info = kern(NUTSNoise(*(x.to(DEV) for x in noise)), s, model, eps, gj.MongeMetric(1.0), M,
                        stopping_criterion=crit)
        so1, oi = okern(noise, so, Mo.logp, eps, metric, max_num_doublings=M, stopping_criterion=crit)
        torch.cuda.synchronize()
        steps = [o.num_integration_steps for o in oi]
        print(f"  t={t}: steps {steps}, turning {[o.is_turning for o in oi]}", flush=True)
        assert info.num_integration_steps.cpu().tolist() == steps
        assert info.num_trajectory_expansions.cpu().tolist() == [o.num_trajectory_expansions for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        assert rel_err(s1.position, so1.position) < 1e-10
        assert rel_err(s1.logdensity, so1.logdensity) < 1e-10
        assert rel_err(s1.logdensity_grad, so1.logdensity_grad) < 1e-10
        assert rel_err(info.energy, torch.tensor([o.energy for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        left = torch.cat([o.trajectory_leftmost_state.position for o in oi])
        right = torch.cat([o.trajectory_rightmost_state.position for o in oi])
        for c, o in enumerate(oi):
            el = rel_err(info.trajectory_leftmost_state[c], left[c])
            er = rel_err(info.trajectory_rightmost_state[c], right[c])
            print(f"    chain {c}: divergent {o.is_divergent}, ends {el:.1e} {er:.1e}", flush=True)
            # The last state of a divergent trajectory is where the energy error
            # passed the threshold: the fixed-point iterates that produced it
            # run away, which amplifies rounding differences.
            bar = 1e-3 if o.is_divergent else 1e-10
            assert el < bar and er < bar, c
        assert rel_err(info.momentum, torch.stack([o.momentum for o in oi])) < 1e-10
        so = so1
        s = RMHMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV))


def test_nuts_large_checkpoint_tables_launch():
    """M = 12 doublings with four checkpoint arrays at four coordinates per lane
    need 96 KiB of LDS per chain (past the 64 KiB default): the launch raises the
    kernel's limit and runs; a short step size keeps every trajectory finite."""
    D, C = 200, 3
    model = gj.NealFunnel(D)
    q = (0.3 * seeded_normal(90, C, D)).to(DEV)
    s = gj.rmhmc.init(q, model)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=8))
    for crit in ("bet", "riem", "euc"):
        s1, info = kern(gr.key(91), s, model, 0.01, gj.MongeMetric(1.0), 12, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"criterion {crit}: steps {info.num_integration_steps.tolist()}", flush=True)
        assert torch.isfinite(s1.position).all()
        assert (info.num_integration_steps >= 1).all() and (info.num_integration_steps <= 2 ** 12 - 1).all()


def test_nuts_keeps_the_funnel_stationary():
    D, C = 10, 4096
    g = torch.Generator().manual_seed(77)
    v = 3.0 * torch.randn(C, 1, generator=g, dtype=torch.float64)
    x = torch.randn(C, D - 1, generator=g, dtype=torch.float64) * torch.exp(0.5 * v)
    q = torch.cat([x, v], dim=1).to(DEV)
    model = gj.NealFunnel(D)
    alg = gj.nuts(model, 0.2, gj.MongeMetric(1.0), max_num_doublings=6,
                  integrator=functools.partial(implicit_midpoint, max_iters=8))
    s = alg.init(q)
    accs, steps = [], []
    for t in range(10):
        s, info = alg.step(gr.key(500 + t), s)
        accs.append(float(info.acceptance_rate.mean()))
        steps.append(float(info.num_integration_steps.double().mean()))
    torch.cuda.synchronize()
    vv = s.position[:, -1].double().cpu()
    m, var = float(vv.mean()), float(vv.var())
    print(f"v mean {m:.3f} (0), var {var:.3f} (9), acceptance {sum(accs) / 10:.3f}, mean steps {sum(steps) / 10:.1f}")
    assert torch.isfinite(s.position).all()
    se_m, se_v = 3.0 / C ** 0.5, 9.0 * (2.0 / C) ** 0.5
    assert abs(m) < 5 * se_m
    assert abs(var - 9.0) < 5 * se_v


@pytest.mark.parametrize("N,D,C,eps,M,T,seed,crit", [
    (120, 6, 3, 0.1, 4, 2, 10, "euc"),
    (120, 6, 3, 0.1, 4, 2, 10, "riem"),
    (256, 64, 2, 0.05, 3, 1, 11, "bet"),   # the kernel's full feature capacity
])
def test_logistic_nuts_matches_oracle(N, D, C, eps, M, T, seed, crit):
    from geomjax_amd.synthetic import make_logistic

    print(f"case logistic N={N} D={D} C={C} eps={eps} M={M} T={T} criterion={crit}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    model = gj.LogisticRegression(X, y, 100.0)
    Mo = om.LogisticRegression(X, y, 100.0)
    q0 = 0.3 * seeded_normal(100 + seed, C, D)
    so = orm.init(q0, Mo.logp)
    s = gj.rmhmc.init(q0.to(DEV), model)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=6))
    okern = onuts.build_kernel(max_iters=6)
    for t in range(T):
        noise = (seeded_normal(200 + 10 * seed + t, C, D), seeded_uniform(300 + 10 * seed + t, C, M),
                 seeded_uniform(400 + 10 * seed + t, C, M), seeded_uniform(500 + 10 * seed + t, C, 2 ** M - 1))
        s1, info = kern(NUTSNoise(*(x.to(DEV) for x in noise)), s, model, eps, gj.FisherMetric(), M,
                        stopping_criterion=crit)
        so1, oi = okern(noise, so, Mo.logp, eps, Mo.fisher_metric_fn, max_num_doublings=M, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {[o.num_integration_steps for o in oi]}, turning {[o.is_turning for o in oi]}",
              flush=True)
        assert info.num_integration_steps.cpu().tolist() == [o.num_integration_steps for o in oi]
        assert info.num_trajectory_expansions.cpu().tolist() == [o.num_trajectory_expansions for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        assert rel_err(s1.position, so1.position) < 1e-10
        assert rel_err(s1.logdensity, so1.logdensity) < 1e-10
        assert rel_err(s1.logdensity_grad, so1.logdensity_grad) < 1e-10
        assert rel_err(info.energy, torch.tensor([o.energy for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.momentum, torch.stack([o.momentum for o in oi])) < 1e-10
        left = torch.cat([o.trajectory_leftmost_state.position for o in oi])
        right = torch.cat([o.trajectory_rightmost_state.position for o in oi])
        assert rel_err(info.trajectory_leftmost_state, left) < 1e-10
        assert rel_err(info.trajectory_rightmost_state, right) < 1e-10
        so = so1
        s = RMHMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV))


def test_logistic_nuts_batch_equals_each_posterior_alone():
    from geomjax_amd.synthetic import make_logistic

    Xa, ya, _ = make_logistic(150, 5, 21)
    Xb, yb, _ = make_logistic(90, 9, 22)
    ma, mb = gj.LogisticRegression(Xa, ya, 100.0), gj.LogisticRegression(Xb, yb, 100.0)
    batch = gj.LogisticBatch([ma, mb], [3, 2])
    C, Dm, M = 5, 9, 5
    q = torch.zeros(C, Dm, dtype=torch.float64)
    q[:3, :5] = 0.3 * seeded_normal(31, 3, 5)
    q[3:, :9] = 0.3 * seeded_normal(32, 2, 9)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    z[:3, :5] = seeded_normal(33, 3, 5)
    z[3:, :9] = seeded_normal(34, 2, 9)
    ud, um, us = seeded_uniform(35, C, M), seeded_uniform(36, C, M), seeded_uniform(37, C, 2 ** M - 1)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=6))
    sb = gj.rmhmc.init(q.to(DEV), batch)
    nb, ib = kern(NUTSNoise(z.to(DEV), ud.to(DEV), um.to(DEV), us.to(DEV)), sb, batch, 0.1, gj.FisherMetric(), M)
    for model, rows, D in ((ma, slice(0, 3), 5), (mb, slice(3, 5), 9)):
        s = gj.rmhmc.init(q[rows, :D].contiguous().to(DEV), model)
        n1, i1 = kern(NUTSNoise(*(x[rows].contiguous().to(DEV) for x in (z[:, :D], ud, um, us))), s, model, 0.1,
                      gj.FisherMetric(), M)
        torch.cuda.synchronize()
        assert torch.equal(ib.num_integration_steps[rows], i1.num_integration_steps)
        assert rel_err(nb.position[rows, :D], n1.position) <= 1e-12
        assert rel_err(nb.logdensity_grad[rows, :D], n1.logdensity_grad) <= 1e-12
        assert rel_err(ib.acceptance_rate[rows], i1.acceptance_rate) <= 1e-12
        assert torch.count_nonzero(nb.position[rows, D:]) == 0  # row tails past the posterior's D


@pytest.mark.parametrize("D,C,eps,M,T,a2,unit_mass,faithful,crit,seed", [
    (10, 4, 0.1, 5, 2, 1.0, True, True, "euc", 20),
    (10, 4, 0.1, 5, 2, 0.6, False, False, "riem", 21),
    (100, 2, 0.05, 4, 1, 1.0, True, True, "bet", 22),
    (100, 2, 0.05, 4, 1, 1.0, True, True, "euc", 23),
])
def test_lmcmonge_nuts_matches_oracle(D, C, eps, M, T, a2, unit_mass, faithful, crit, seed):
    from geomjax_amd.lmcmonge import LMCState
    from geomjax_amd.lmcmonge import init as lmc_init
    from geomjax_amd.lmcmonge_nuts import build_kernel as lnuts_kernel
    from oracle import lmcmonge as olm

    print(f"case lmcmonge D={D} C={C} eps={eps} M={M} T={T} alpha2={a2} unit_mass={unit_mass} faithful={faithful} "
          f"criterion={crit}", flush=True)
    Mo = om.NealFunnel(D)
    model = gj.NealFunnel(D)
    im = torch.ones(D, dtype=torch.float64) if unit_mass else 0.5 + seeded_uniform(600 + seed, D)
    q0 = 0.5 * seeded_normal(610 + seed, C, D)
    so = olm.init(q0, Mo.logp)
    s = lmc_init(q0.to(DEV), model)
    kern = lnuts_kernel(max_num_doublings=M, reference_faithful=faithful)
    okern = onuts.build_lmcmonge_kernel(max_num_doublings=M, reference_faithful=faithful)
    for t in range(T):
        noise = (seeded_normal(620 + 10 * seed + t, C, D), seeded_uniform(630 + 10 * seed + t, C, M),
                 seeded_uniform(640 + 10 * seed + t, C, M), seeded_uniform(650 + 10 * seed + t, C, 2 ** M - 1))
        s1, info = kern(NUTSNoise(*(x.to(DEV) for x in noise)), s, model, eps, im.to(DEV), a2,
                        stopping_criterion=crit)
        so1, oi = okern(noise, so, Mo.logp, eps, im, a2, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {[o.num_integration_steps for o in oi]}, turning {[o.is_turning for o in oi]}, "
              f"divergent {[o.is_divergent for o in oi]}", flush=True)
        assert info.num_integration_steps.cpu().tolist() == [o.num_integration_steps for o in oi]
        assert info.num_trajectory_expansions.cpu().tolist() == [o.num_trajectory_expansions for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        assert rel_err(s1.position, so1.position) < 1e-10
        assert rel_err(s1.logdensity, so1.logdensity) < 1e-10
        assert rel_err(s1.logdensity_grad, so1.logdensity_grad) < 1e-10
        jscale = max(1.0, so1.volume_adjustment.abs().max().item())
        assert (s1.volume_adjustment.cpu() - so1.volume_adjustment).abs().max().item() < 1e-10 * jscale
        assert rel_err(info.energy, torch.tensor([o.energy for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.velocity, torch.stack([o.momentum for o in oi])) < 1e-10
        for c, o in enumerate(oi):
            bar = 1e-3 if o.is_divergent else 1e-10
            assert rel_err(info.trajectory_leftmost_state[c], o.trajectory_leftmost_state.position[0]) < bar, c
            assert rel_err(info.trajectory_rightmost_state[c], o.trajectory_rightmost_state.position[0]) < bar, c
        so = so1
        s = LMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV),
                     so.volume_adjustment.to(DEV))


def test_lmcmonge_nuts_keeps_the_funnel_stationary():
    """Exact algebra (reference_faithful=False): chains started from exact funnel
    draws keep v's mean and variance after ten transitions (5 standard errors)."""
    D, C = 10, 4096
    g = torch.Generator().manual_seed(78)
    v = 3.0 * torch.randn(C, 1, generator=g, dtype=torch.float64)
    x = torch.randn(C, D - 1, generator=g, dtype=torch.float64) * torch.exp(0.5 * v)
    q = torch.cat([x, v], dim=1).to(DEV)
    model = gj.NealFunnel(D)
    im = torch.ones(D, dtype=torch.float64, device=DEV)
    alg = gj.nutslmcmonge(model, 0.1, im, alpha2=1.0, max_num_doublings=6, reference_faithful=False)
    s = alg.init(q)
    accs = []
    for t in range(10):
        s, info = alg.step(gr.key(700 + t), s)
        accs.append(float(info.acceptance_rate.mean()))
    torch.cuda.synchronize()
    vv = s.position[:, -1].cpu()
    m, var = float(vv.mean()), float(vv.var())
    print(f"v mean {m:.3f} (0), var {var:.3f} (9), acceptance {sum(accs) / 10:.3f}")
    assert torch.isfinite(s.position).all()
    assert abs(m) < 5 * 3.0 / C ** 0.5
    assert abs(var - 9.0) < 5 * 9.0 * (2.0 / C) ** 0.5


@pytest.mark.parametrize("N,D,C,eps,M,T,seed,crit", [
    (120, 6, 3, 0.6, 4, 2, 40, "euc"),
    (120, 6, 3, 0.6, 4, 2, 41, "riem"),
    (200, 12, 3, 0.5, 5, 2, 42, "bet"),
    (120, 6, 3, 3.0, 4, 2, 44, "euc"),     # divergent trajectories
    (256, 64, 2, 0.05, 3, 1, 43, "euc"),   # the kernel's full feature capacity
])
def test_lmcmc_nuts_matches_oracle(N, D, C, eps, M, T, seed, crit):
    from geomjax_amd.lmcmc import LMCState
    from geomjax_amd.lmcmc import init as lmc_init
    from geomjax_amd.lmcmc_nuts import build_kernel as lnuts_kernel
    from geomjax_amd.synthetic import make_logistic
    from oracle import lmcmc as olc

    print(f"case lmcmc-nuts N={N} D={D} C={C} eps={eps} M={M} T={T} criterion={crit}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    model = gj.LogisticRegression(X, y, 100.0)
    Mo = om.LogisticRegression(X, y, 100.0)
    q0 = 0.3 * seeded_normal(700 + seed, C, D)
    so = olc.init(q0, Mo.logp)
    s = lmc_init(q0.to(DEV), model)
    kern = lnuts_kernel()
    okern = onuts.build_lmcmc_kernel()
    for t in range(T):
        noise = (seeded_normal(720 + 10 * seed + t, C, D), seeded_uniform(730 + 10 * seed + t, C, M),
                 seeded_uniform(740 + 10 * seed + t, C, M), seeded_uniform(750 + 10 * seed + t, C, 2 ** M - 1))
        s1, info = kern(NUTSNoise(*(x.to(DEV) for x in noise)), s, model, eps, gj.FisherMetric(), M,
                        stopping_criterion=crit)
        so1, oi = okern(noise, so, Mo.logp, eps, Mo.fisher_metric_fn, max_num_doublings=M, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {[o.num_integration_steps for o in oi]}, turning {[o.is_turning for o in oi]}, "
              f"divergent {[o.is_divergent for o in oi]}", flush=True)
        assert info.num_integration_steps.cpu().tolist() == [o.num_integration_steps for o in oi]
        assert info.num_trajectory_expansions.cpu().tolist() == [o.num_trajectory_expansions for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        assert rel_err(s1.position, so1.position) < 1e-10
        assert rel_err(s1.logdensity, so1.logdensity) < 1e-10
        assert rel_err(s1.logdensity_grad, so1.logdensity_grad) < 1e-10
        jscale = max(1.0, so1.volume_adjustment.abs().max().item())
        assert (s1.volume_adjustment.cpu() - so1.volume_adjustment).abs().max().item() < 1e-10 * jscale
        assert rel_err(info.energy, torch.tensor([o.energy for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.velocity, torch.stack([o.momentum for o in oi])) < 1e-10
        for c, o in enumerate(oi):
            bar = 1e-3 if o.is_divergent else 1e-10
            assert rel_err(info.trajectory_leftmost_state[c], o.trajectory_leftmost_state.position[0]) < bar, c
            assert rel_err(info.trajectory_rightmost_state[c], o.trajectory_rightmost_state.position[0]) < bar, c
        so = so1
        s = LMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV),
                     so.volume_adjustment.to(DEV))


def test_lmcmc_nuts_batch_equals_each_posterior_alone():
    from geomjax_amd.lmcmc import init as lmc_init
    from geomjax_amd.lmcmc_nuts import build_kernel as lnuts_kernel
    from geomjax_amd.synthetic import make_logistic

    Xa, ya, _ = make_logistic(150, 5, 51)
    Xb, yb, _ = make_logistic(90, 9, 52)
    ma, mb = gj.LogisticRegression(Xa, ya, 100.0), gj.LogisticRegression(Xb, yb, 100.0)
    batch = gj.LogisticBatch([ma, mb], [3, 2])
    C, Dm, M = 5, 9, 5
    q = torch.zeros(C, Dm, dtype=torch.float64)
    q[:3, :5] = 0.3 * seeded_normal(61, 3, 5)
    q[3:, :9] = 0.3 * seeded_normal(62, 2, 9)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    z[:3, :5] = seeded_normal(63, 3, 5)
    z[3:, :9] = seeded_normal(64, 2, 9)
    ud, um, us = seeded_uniform(65, C, M), seeded_uniform(66, C, M), seeded_uniform(67, C, 2 ** M - 1)
    kern = lnuts_kernel()
    sb = lmc_init(q.to(DEV), batch)
    nb, ib = kern(NUTSNoise(z.to(DEV), ud.to(DEV), um.to(DEV), us.to(DEV)), sb, batch, 0.2, gj.FisherMetric(), M)
    for model, rows, D in ((ma, slice(0, 3), 5), (mb, slice(3, 5), 9)):
        s = lmc_init(q[rows, :D].contiguous().to(DEV), model)
        n1, i1 = kern(NUTSNoise(*(x[rows].contiguous().to(DEV) for x in (z[:, :D], ud, um, us))), s, model, 0.2,
                      gj.FisherMetric(), M)
        torch.cuda.synchronize()
        assert torch.equal(ib.num_integration_steps[rows], i1.num_integration_steps)
        assert rel_err(nb.position[rows, :D], n1.position) <= 1e-12
        assert rel_err(nb.volume_adjustment[rows], n1.volume_adjustment) <= 1e-12
        assert rel_err(ib.acceptance_rate[rows], i1.acceptance_rate) <= 1e-12
        assert torch.count_nonzero(nb.position[rows, D:]) == 0  # row tails past the posterior's D


def test_nutslmc_sampling_algorithm_runs_on_device_noise():
    """gj.nutslmc's SamplingAlgorithm with a device Philox key: finite states,
    step counts within 2^M - 1, acceptance in [0, 1]."""
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(300, 10, 71)
    model = gj.LogisticRegression(X, y, 100.0)
    alg = gj.nutslmc(model, 0.2, gj.FisherMetric(), max_num_doublings=6)
    s = alg.init(0.1 * seeded_normal(72, 256, 10).to(DEV))
    for t in range(3):
        s, info = alg.step(gr.key(800 + t), s)
    torch.cuda.synchronize()
    assert torch.isfinite(s.position).all() and torch.isfinite(s.volume_adjustment).all()
    steps = info.num_integration_steps.cpu()
    assert (steps >= 1).all() and (steps <= 2 ** 6 - 1).all()
    acc = info.acceptance_rate.cpu()
    assert ((acc >= 0) & (acc <= 1)).all()


@pytest.mark.parametrize("crit", ["euc", "riem"])
def test_persistent_nuts_launch_equals_step_launches(crit):
    """T transitions in one launch (the chains keep their state on chip) equal T
    single-transition launches fed the same draws: host-fed NUTSNoise with a
    leading [T] axis, and the device PRNG (transition t at counter + t)."""
    from geomjax_amd.nuts import NUTSNoise, sample

    D, C, eps, M, T = 12, 6, 0.1, 5, 4
    print(f"case persistent NUTS D={D} C={C} M={M} T={T} criterion={crit}", flush=True)
    model = gj.NealFunnel(D)
    integ = functools.partial(implicit_midpoint, max_iters=8)
    kern = nuts_kernel(integ)
    s0 = gj.rmhmc.init((0.5 * seeded_normal(140, C, D)).to(DEV), model)
    z = seeded_normal(141, T, C, D).to(DEV)
    ud, um = seeded_uniform(142, T, C, M).to(DEV), seeded_uniform(143, T, C, M).to(DEV)
    us = seeded_uniform(144, T, C, 2 ** M - 1).to(DEV)
    res = sample(NUTSNoise(z, ud, um, us), s0, model, eps, gj.MongeMetric(1.0), T, max_num_doublings=M,
                 stopping_criterion=crit, integrator=integ)
    s = s0
    for t in range(T):
        s, info = kern(NUTSNoise(z[t], ud[t], um[t], us[t]), s, model, eps, gj.MongeMetric(1.0), M,
                       stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {info.num_integration_steps.tolist()} vs {res.num_integration_steps[t].tolist()}, "
              f"max |position difference| {(res.samples[t] - s.position).abs().max().item():.3e}", flush=True)
        assert torch.equal(res.num_integration_steps[t], info.num_integration_steps)
        assert rel_err(res.samples[t], s.position) < 1e-13
        assert torch.allclose(res.acceptance_rate[t].double(), info.acceptance_rate, rtol=1e-6)
    torch.cuda.synchronize()
    assert rel_err(res.state.position, s.position) < 1e-13
    assert rel_err(res.state.logdensity, s.logdensity) < 1e-13
    assert rel_err(res.state.logdensity_grad, s.logdensity_grad) < 1e-13
    assert torch.equal(res.num_fevals, res.num_integration_steps.to(torch.int64).sum(0))
    # device PRNG
    key = gr.PRNGKey(7, 100)
    res = sample(key, s0, model, eps, gj.MongeMetric(1.0), T, max_num_doublings=M, stopping_criterion=crit,
                 integrator=integ)
    s = s0
    for t in range(T):
        s, _ = kern(gr.PRNGKey(7, 100 + t), s, model, eps, gj.MongeMetric(1.0), M, stopping_criterion=crit)
    torch.cuda.synchronize()
    assert rel_err(res.state.position, s.position) < 1e-13


# ---------------------------------------------------------------------------
# Tuned (float32) form of NUTS-RMHMC on the funnel
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("D,C,eps,M", [(10, 64, 0.1, 6), (100, 64, 0.03, 8)])
def test_tuned_nuts_float32_tracks_float64(D, C, eps, M):
    """The float32 form on the same host draws as the float64 kernel: the tree
    decisions (uniforms against float64 weights, U-turn signs of float64 dot
    products) follow the same path on nearly every chain, so the step counts
    agree on at least 90 % of the chains and, on those, the new positions agree
    to 1e-3 relative (float32 trajectories: every fixed-point iterate rounds to
    float32)."""
    print(f"case tuned NUTS D={D} C={C} eps={eps} M={M}", flush=True)
    model = gj.NealFunnel(D)
    integ = functools.partial(implicit_midpoint, max_iters=8)
    kern = nuts_kernel(integ)
    q0 = 0.5 * seeded_normal(150, C, D)
    z = seeded_normal(151, C, D)
    ud, um, us = seeded_uniform(152, C, M), seeded_uniform(153, C, M), seeded_uniform(154, C, 2 ** M - 1)
    noise = lambda dt: NUTSNoise(z.to(DEV, dt), ud.to(DEV), um.to(DEV), us.to(DEV))
    s64, i64 = kern(noise(torch.float64), gj.rmhmc.init(q0.to(DEV), model), model, eps, gj.MongeMetric(1.0), M)
    s32, i32 = kern(noise(torch.float32), gj.rmhmc.init(q0.to(DEV).float(), model), model, eps, gj.MongeMetric(1.0), M)
    torch.cuda.synchronize()
    assert s32.position.dtype == torch.float32 and s32.logdensity.dtype == torch.float64
    same = (i32.num_integration_steps == i64.num_integration_steps).cpu()
    frac = float(same.double().mean())
    err = rel_err(s32.position[same.to(DEV)].double(), s64.position[same.to(DEV)])
    print(f"  same step counts on {frac:.3f} of the chains; position error there {err:.2e}", flush=True)
    assert torch.isfinite(s32.position).all()
    assert frac >= 0.9
    assert err < 1e-3


def test_tuned_nuts_keeps_the_funnel_stationary():
    """Exact funnel draws stay exact under ten float32 NUTS transitions (device
    PRNG): v's mean and variance within 5 standard errors, as for float64."""
    D, C = 10, 4096
    g = torch.Generator().manual_seed(78)
    v = 3.0 * torch.randn(C, 1, generator=g, dtype=torch.float64)
    x = torch.randn(C, D - 1, generator=g, dtype=torch.float64) * torch.exp(0.5 * v)
    q = torch.cat([x, v], dim=1).to(DEV).float()
    model = gj.NealFunnel(D)
    alg = gj.nuts(model, 0.2, gj.MongeMetric(1.0), max_num_doublings=6,
                  integrator=functools.partial(implicit_midpoint, max_iters=8))
    s = alg.init(q)
    for t in range(10):
        s, info = alg.step(gr.key(600 + t), s)
    torch.cuda.synchronize()
    vv = s.position[:, -1].double().cpu()
    m, var = float(vv.mean()), float(vv.var())
    print(f"float32: v mean {m:.3f} (0), var {var:.3f} (9)")
    assert torch.isfinite(s.position).all()
    se_m, se_v = 3.0 / C ** 0.5, 9.0 * (2.0 / C) ** 0.5
    assert abs(m) < 5 * se_m
    assert abs(var - 9.0) < 5 * se_v


def test_tuned_persistent_nuts_equals_step_launches():
    """The float32 persistent launch equals float32 single-transition launches
    on the same host draws: step counts identical, the first transition
    bit-identical, later states within the float32-vs-float64 bar (1e-3).
    The two forms carry the chain between transitions in registers and in
    memory; float32 rounding differences of the same code then grow along
    the funnel's neck (the float64 form of this check agrees to 1e-13)."""
    from geomjax_amd.nuts import sample

    D, C, eps, M, T = 12, 6, 0.1, 5, 4
    model = gj.NealFunnel(D)
    integ = functools.partial(implicit_midpoint, max_iters=8)
    kern = nuts_kernel(integ)
    s0 = gj.rmhmc.init((0.5 * seeded_normal(160, C, D)).to(DEV).float(), model)
    z = seeded_normal(161, T, C, D).to(DEV).float()
    ud, um = seeded_uniform(162, T, C, M).to(DEV), seeded_uniform(163, T, C, M).to(DEV)
    us = seeded_uniform(164, T, C, 2 ** M - 1).to(DEV)
    res = sample(NUTSNoise(z, ud, um, us), s0, model, eps, gj.MongeMetric(1.0), T, max_num_doublings=M,
                 integrator=integ)
    s = s0
    for t in range(T):
        s, info = kern(NUTSNoise(z[t], ud[t], um[t], us[t]), s, model, eps, gj.MongeMetric(1.0), M)
        torch.cuda.synchronize()
        assert torch.equal(res.num_integration_steps[t], info.num_integration_steps)
        err = rel_err(res.samples[t].double(), s.position.double())
        print(f"tuned persistent NUTS transition {t}: rel err {err:.2e}")
        if t == 0:
            assert torch.equal(res.samples[0], s.position)
        assert err < 1e-3
    assert res.samples.dtype == torch.float32


@pytest.mark.parametrize("N,D,C,eps,M", [(256, 8, 32, 0.1, 5), (1024, 40, 32, 0.05, 5), (2048, 64, 16, 0.05, 4)])
def test_tuned_logistic_nuts_tracks_float64(N, D, C, eps, M):
    """The tuned logistic form (float32 state, split-bf16 MFMA, float64 energies
    and tree weights) on the same host draws as the float64 kernel, which
    equals the oracle at 1e-10: on at least 90 % of the chains the step count is
    the same and the new position agrees to 1e-3 relative (a float32 weight can
    flip a progressive-sampling or U-turn decision on the odd chain)."""
    from geomjax_amd.synthetic import make_logistic

    print(f"case tuned logistic NUTS N={N} D={D} C={C} eps={eps} M={M}", flush=True)
    X, y, _ = make_logistic(N, D, 30 + D)
    model = gj.LogisticRegression(X, y, 100.0)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=6))
    q0 = (0.3 * seeded_normal(160 + D, C, D)).float().double()
    z = seeded_normal(161 + D, C, D).float().double()
    ud, um, us = seeded_uniform(162, C, M), seeded_uniform(163, C, M), seeded_uniform(164, C, 2 ** M - 1)
    noise = lambda dt: NUTSNoise(z.to(DEV, dt), ud.to(DEV), um.to(DEV), us.to(DEV))
    s64, i64 = kern(noise(torch.float64), gj.rmhmc.init(q0.to(DEV), model), model, eps, gj.FisherMetric(), M)
    s32, i32 = kern(noise(torch.float32), gj.rmhmc.init(q0.to(DEV).float(), model), model, eps, gj.FisherMetric(), M)
    torch.cuda.synchronize()
    assert s32.position.dtype == torch.float32 and s32.logdensity.dtype == torch.float64
    same = (i32.num_integration_steps == i64.num_integration_steps).cpu()
    per = ((s32.position.double() - s64.position).norm(dim=1) / s64.position.norm(dim=1).clamp_min(1e-300)).cpu()
    close = same & (per < 1e-3)
    frac = float(close.double().mean())
    print(f"  steps {i64.num_integration_steps.tolist()}", flush=True)
    print(f"  same steps on {float(same.double().mean()):.3f}, same steps and position within 1e-3 on {frac:.3f}; "
          f"median position error {float(per[same].median()) if bool(same.any()) else float('nan'):.1e}; "
          f"energy diff there {(i32.energy - i64.energy).abs().cpu()[close].max().item():.1e}", flush=True)
    assert torch.isfinite(s32.position).all()
    assert frac >= 0.9
    assert (i32.energy - i64.energy).abs().cpu()[close].max().item() < 1e-2


def test_tuned_logistic_nuts_batch_equals_each_posterior_alone():
    from geomjax_amd.synthetic import make_logistic

    shapes = [(300, 8), (400, 40)]
    models = [gj.LogisticRegression(*make_logistic(n, d, 50 + i)[:2], 100.0) for i, (n, d) in enumerate(shapes)]
    cpp, M = 4, 4
    batch = gj.LogisticBatch(models, chains_per_posterior=cpp)
    C = cpp * len(models)
    q = torch.zeros(C, batch.dim)
    z = torch.zeros(C, batch.dim)
    for i, (_, d) in enumerate(shapes):
        q[i * cpp:(i + 1) * cpp, :d] = 0.3 * seeded_normal(170 + i, cpp, d)
        z[i * cpp:(i + 1) * cpp, :d] = seeded_normal(175 + i, cpp, d)
    ud, um, us = seeded_uniform(180, C, M), seeded_uniform(181, C, M), seeded_uniform(182, C, 2 ** M - 1)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=6))
    sb = gj.rmhmc.init(q.float().to(DEV), batch)
    sb1, ib = kern(NUTSNoise(z.float().to(DEV), ud.to(DEV), um.to(DEV), us.to(DEV)), sb, batch, 0.1,
                   gj.FisherMetric(), M)
    for i, (_, d) in enumerate(shapes):
        rows = slice(i * cpp, (i + 1) * cpp)
        s = gj.rmhmc.init(q[rows, :d].float().to(DEV), models[i])
        s1, info = kern(NUTSNoise(z[rows, :d].float().to(DEV), ud[rows].to(DEV), um[rows].to(DEV), us[rows].to(DEV)),
                        s, models[i], 0.1, gj.FisherMetric(), M)
        torch.cuda.synchronize()
        err = rel_err(sb1.position[rows, :d].double(), s1.position.double())
        print(f"posterior {i} (D={d}): steps {info.num_integration_steps.tolist()}, rel err {err:.1e}", flush=True)
        assert torch.equal(ib.num_integration_steps[rows], info.num_integration_steps)
        assert err < 1e-6
        assert torch.all(sb1.position[rows, d:] == 0)


def test_tuned_logistic_nuts_chains_track_float64_chains():
    """Paired with the float64 kernel on the device PRNG over ten transitions:
    chain means within 5 standard errors, mean acceptance statistics within
    0.02 and mean step counts within 5 %."""
    from geomjax_amd.synthetic import make_logistic

    N, D, C, eps, M, T = 512, 8, 1024, 0.15, 5, 10
    X, y, _ = make_logistic(N, D, 61)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = 0.1 * seeded_normal(190, C, D)
    alg = gj.nuts(model, eps, gj.FisherMetric(), max_num_doublings=M,
                  integrator=functools.partial(implicit_midpoint, max_iters=6))
    out = {}
    for dt in (torch.float64, torch.float32):
        s = alg.init(q0.to(DEV).to(dt))
        accs, steps = [], []
        for t in range(T):
            s, info = alg.step(gr.key(700 + t), s)
            accs.append(float(info.acceptance_rate.mean()))
            steps.append(float(info.num_integration_steps.double().mean()))
        torch.cuda.synchronize()
        out[dt] = (s.position.double().cpu(), sum(accs) / T, sum(steps) / T)
    (p64, a64, n64), (p32, a32, n32) = out[torch.float64], out[torch.float32]
    se = p64.std(0) / C ** 0.5
    dm = (p32.mean(0) - p64.mean(0)).abs()
    print(f"acceptance {a64:.4f} vs {a32:.4f}; steps {n64:.2f} vs {n32:.2f}; mean diff / SE max "
          f"{(dm / se).max().item():.2f}", flush=True)
    assert abs(a64 - a32) < 0.02
    assert abs(n64 - n32) < 0.05 * n64
    assert torch.all(dm < 5 * se)


@pytest.mark.parametrize("D,C,eps,M,faithful", [(10, 64, 0.1, 6, True), (100, 64, 0.03, 7, False)])
def test_tuned_lmcmonge_nuts_tracks_float64(D, C, eps, M, faithful):
    """The float32 LMC-Monge NUTS form (float32 trajectory, float64 log density,
    kinetic-energy sums, volume change and tree weights) on the same host draws
    as the float64 kernel, in both algebras: on at least 90 % of the chains the
    step count is the same and the new position and volume adjustment agree to
    1e-3 (relative / absolute)."""
    from geomjax_amd.lmcmonge_nuts import build_kernel as lnuts_kernel

    print(f"case tuned LMC-Monge NUTS D={D} C={C} eps={eps} M={M} faithful={faithful}", flush=True)
    model = gj.NealFunnel(D)
    kern = lnuts_kernel(max_num_doublings=M, reference_faithful=faithful)
    q0 = (0.5 * seeded_normal(200, C, D)).float().double()
    z = seeded_normal(201, C, D).float().double()
    ud, um, us = seeded_uniform(202, C, M), seeded_uniform(203, C, M), seeded_uniform(204, C, 2 ** M - 1)
    noise = lambda dt: NUTSNoise(z.to(DEV, dt), ud.to(DEV), um.to(DEV), us.to(DEV))
    out = {}
    for dt in (torch.float64, torch.float32):
        im = torch.ones(D, dtype=dt, device=DEV)
        out[dt] = kern(noise(dt), gj.lmcmonge.init(q0.to(DEV, dt), model), model, eps, im, 1.0)
    torch.cuda.synchronize()
    (s64, i64), (s32, i32) = out[torch.float64], out[torch.float32]
    assert s32.position.dtype == torch.float32 and s32.logdensity.dtype == torch.float64
    same = (i32.num_integration_steps == i64.num_integration_steps).cpu()
    per = ((s32.position.double() - s64.position).norm(dim=1) / s64.position.norm(dim=1)).cpu()
    dj = (s32.volume_adjustment - s64.volume_adjustment).abs().cpu()
    close = same & (per < 1e-3) & (dj < 1e-3)
    frac = float(close.double().mean())
    print(f"  same steps on {float(same.double().mean()):.3f}, close on {frac:.3f}; median position error "
          f"{float(per[same].median()):.1e}, max J diff there {float(dj[close].max()):.1e}", flush=True)
    assert torch.isfinite(s32.position).all()
    assert frac >= 0.9


def test_tuned_lmcmonge_nuts_keeps_the_funnel_stationary():
    """Exact algebra, float32 state: exact funnel draws keep v's mean and
    variance over ten transitions (5 standard errors), as the float64 form."""
    D, C = 10, 4096
    g = torch.Generator().manual_seed(79)
    v = 3.0 * torch.randn(C, 1, generator=g, dtype=torch.float64)
    x = torch.randn(C, D - 1, generator=g, dtype=torch.float64) * torch.exp(0.5 * v)
    q = torch.cat([x, v], dim=1).to(DEV).float()
    model = gj.NealFunnel(D)
    im = torch.ones(D, dtype=torch.float32, device=DEV)
    alg = gj.nutslmcmonge(model, 0.1, im, alpha2=1.0, max_num_doublings=6, reference_faithful=False)
    s = alg.init(q)
    accs = []
    for t in range(10):
        s, info = alg.step(gr.key(600 + t), s)
        accs.append(float(info.acceptance_rate.mean()))
    torch.cuda.synchronize()
    vv = s.position[:, -1].double().cpu()
    m, var = float(vv.mean()), float(vv.var())
    print(f"v mean {m:.3f} (0), var {var:.3f} (9), acceptance {sum(accs) / 10:.3f}", flush=True)
    assert s.position.dtype == torch.float32 and torch.isfinite(s.position).all()
    se_m, se_v = 3.0 / C ** 0.5, 9.0 * (2.0 / C) ** 0.5
    assert abs(m) < 5 * se_m
    assert abs(var - 9.0) < 5 * se_v


@pytest.mark.parametrize("N,D,C,eps,M,crit", [(256, 8, 32, 0.3, 5, "euc"), (1024, 40, 32, 0.15, 5, "riem"),
                                              (2048, 64, 16, 0.1, 4, "bet")])
def test_tuned_lmcmc_nuts_tracks_float64(N, D, C, eps, M, crit):
    """The tuned NUTS-LMC form (float32 trajectory, split-bf16 MFMA Gram passes,
    float64 log densities, log-determinants, volume change and weights) on the
    same host draws as the float64 kernel, which equals the oracle at 1e-10: on
    at least 90 % of the chains the step count is the same and the new position
    (1e-3 relative) and volume adjustment (1e-2: four float32 log-determinants
    per integration step, 15 steps here, observed up to ~1e-3) agree."""
    from geomjax_amd.lmcmc import init as lmc_init
    from geomjax_amd.lmcmc_nuts import build_kernel as lnuts_kernel
    from geomjax_amd.synthetic import make_logistic

    print(f"case tuned lmcmc-nuts N={N} D={D} C={C} eps={eps} M={M} criterion={crit}", flush=True)
    X, y, _ = make_logistic(N, D, 80 + D)
    model = gj.LogisticRegression(X, y, 100.0)
    kern = lnuts_kernel()
    q0 = (0.3 * seeded_normal(210 + D, C, D)).float().double()
    z = seeded_normal(211 + D, C, D).float().double()
    ud, um, us = seeded_uniform(212, C, M), seeded_uniform(213, C, M), seeded_uniform(214, C, 2 ** M - 1)
    noise = lambda dt: NUTSNoise(z.to(DEV, dt), ud.to(DEV), um.to(DEV), us.to(DEV))
    out = {dt: kern(noise(dt), lmc_init(q0.to(DEV, dt), model), model, eps, gj.FisherMetric(), M,
                    stopping_criterion=crit) for dt in (torch.float64, torch.float32)}
    torch.cuda.synchronize()
    (s64, i64), (s32, i32) = out[torch.float64], out[torch.float32]
    assert s32.position.dtype == torch.float32 and s32.logdensity.dtype == torch.float64
    same = (i32.num_integration_steps == i64.num_integration_steps).cpu()
    per = ((s32.position.double() - s64.position).norm(dim=1) / s64.position.norm(dim=1)).cpu()
    dj = (s32.volume_adjustment - s64.volume_adjustment).abs().cpu()
    close = same & (per < 1e-3) & (dj < 1e-2)
    frac = float(close.double().mean())
    print(f"  steps {i64.num_integration_steps.tolist()[:8]}; same steps on {float(same.double().mean()):.3f}, "
          f"close on {frac:.3f}; median position error {float(per[same].median()):.1e}, "
          f"max J diff there {float(dj[close].max()) if bool(close.any()) else float('nan'):.1e}", flush=True)
    assert torch.isfinite(s32.position).all()
    assert frac >= 0.9
