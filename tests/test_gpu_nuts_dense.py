"""GPU parity: NUTS with the RMHMC integrator for a user dense metric (the
funnel's Monge metric written as a torch function, differentiated by
torch.func) against the oracle's per-chain expansion (oracle/nuts.py) on the
same draws: step counts, expansions and flags identical; states, energies,
acceptance statistics and trajectory ends 1e-10 (1e-3 for the ends of a
divergent trajectory, whose fixed-point iterates ran away)."""
import functools
import math

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.nuts import NUTSNoise
from geomjax_amd.nuts import build_kernel as nuts_kernel
from geomjax_amd.rmhmc import RMHMCState, implicit_midpoint
from geomjax_amd.rmhmc import init as rm_init
from oracle import nuts as onuts
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def funnel_logp(q):
    x, v = q[:-1], q[-1]
    return (-0.5 * v * v / 9.0 - 0.5 * math.log(2 * math.pi * 9.0)
            + torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi)))


def monge(q):
    g = torch.func.grad(funnel_logp)(q)
    return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + torch.outer(g, g)


@pytest.mark.parametrize("D,C,eps,M,crit,seed", [(6, 3, 0.2, 4, "euc", 1), (6, 3, 0.2, 4, "riem", 2),
                                                (8, 2, 0.3, 3, "bet", 3)])
def test_dense_nuts_matches_oracle(D, C, eps, M, crit, seed):
    print(f"case dense nuts D={D} C={C} eps={eps} M={M} criterion={crit}", flush=True)
    q0 = 0.5 * seeded_normal(10 + seed, C, D)
    so = orm.init(q0, funnel_logp)
    s = rm_init(q0.to(DEV), funnel_logp)
    kern = nuts_kernel(functools.partial(implicit_midpoint, max_iters=8))
    okern = onuts.build_kernel(max_iters=8)
    for t in range(2):
        noise = (seeded_normal(20 + 10 * seed + t, C, D), seeded_uniform(30 + 10 * seed + t, C, M),
                 seeded_uniform(40 + 10 * seed + t, C, M), seeded_uniform(50 + 10 * seed + t, C, 2 ** M - 1))
        s1, info = kern(NUTSNoise(*(x.to(DEV) for x in noise)), s, funnel_logp, eps, monge, M, stopping_criterion=crit)
        so1, oi = okern(noise, so, funnel_logp, eps, monge, max_num_doublings=M, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {[o.num_integration_steps for o in oi]}, turning {[o.is_turning for o in oi]}, "
              f"divergent {[o.is_divergent for o in oi]}", flush=True)
        assert info.num_integration_steps.cpu().tolist() == [o.num_integration_steps for o in oi]
        assert info.num_trajectory_expansions.cpu().tolist() == [o.num_trajectory_expansions for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        for f in ("position", "logdensity", "logdensity_grad"):
            assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, f
        assert rel_err(info.energy, torch.tensor([o.energy for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.momentum, torch.stack([o.momentum for o in oi])) < 1e-10
        for c, o in enumerate(oi):
            bar = 1e-3 if o.is_divergent else 1e-10
            assert rel_err(info.trajectory_leftmost_state[c], o.trajectory_leftmost_state.position[0]) < bar, c
            assert rel_err(info.trajectory_rightmost_state[c], o.trajectory_rightmost_state.position[0]) < bar, c
        so = so1
        s = RMHMCState(*(x.to(DEV) for x in so))


def test_dense_nuts_api_with_device_draws():
    alg = gj.nutsrmhmc(funnel_logp, 0.2, monge, max_num_doublings=4,
                       integrator=functools.partial(implicit_midpoint, max_iters=6))
    s = alg.init(0.5 * seeded_normal(60, 16, 6).to(DEV))
    s, info = alg.step(gr.key(7), s)
    torch.cuda.synchronize()
    assert torch.isfinite(s.position).all()
    steps = info.num_integration_steps.cpu()
    assert (steps >= 1).all() and (steps <= 15).all()


@pytest.mark.parametrize("crit", ["euc", "riem"])
def test_dense_lmc_nuts_matches_oracle(crit):
    """NUTS with the general-metric LMC integrator and a user metric (funnel,
    Monge metric by autodiff) against oracle/nuts.py's lmcmc expansion."""
    from geomjax_amd.lmcmc import LMCState
    from geomjax_amd.lmcmc import init as lmc_init
    from geomjax_amd.lmcmc_nuts import build_kernel as lnuts_kernel
    from oracle import lmcmc as olc

    D, C, eps, M = 6, 3, 0.2, 4
    print(f"case dense lmc-nuts D={D} C={C} eps={eps} M={M} criterion={crit}", flush=True)
    q0 = 0.5 * seeded_normal(70, C, D)
    so = olc.init(q0, funnel_logp)
    s = lmc_init(q0.to(DEV), funnel_logp)
    okern = onuts.build_lmcmc_kernel()
    for t in range(2):
        noise = (seeded_normal(71 + t, C, D), seeded_uniform(73 + t, C, M), seeded_uniform(75 + t, C, M),
                 seeded_uniform(77 + t, C, 2 ** M - 1))
        s1, info = lnuts_kernel()(NUTSNoise(*(x.to(DEV) for x in noise)), s, funnel_logp, eps, monge, M,
                                  stopping_criterion=crit)
        so1, oi = okern(noise, so, funnel_logp, eps, monge, max_num_doublings=M, stopping_criterion=crit)
        torch.cuda.synchronize()
        print(f"  t={t}: steps {[o.num_integration_steps for o in oi]}", flush=True)
        assert info.num_integration_steps.cpu().tolist() == [o.num_integration_steps for o in oi]
        assert info.is_turning.cpu().tolist() == [o.is_turning for o in oi]
        assert info.is_divergent.cpu().tolist() == [o.is_divergent for o in oi]
        for f in ("position", "logdensity", "logdensity_grad"):
            assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, f
        assert (s1.volume_adjustment.cpu() - so1.volume_adjustment).abs().max() < 1e-10 * max(
            1.0, so1.volume_adjustment.abs().max().item())
        assert rel_err(info.acceptance_rate, torch.tensor([o.acceptance_rate for o in oi], dtype=torch.float64)) < 1e-10
        assert rel_err(info.velocity, torch.stack([o.momentum for o in oi])) < 1e-10
        so = so1
        s = LMCState(*(x.to(DEV) for x in so))
