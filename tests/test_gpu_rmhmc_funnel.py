"""GPU parity: RMHMC (implicit midpoint) with the Monge metric on Neal's funnel.

The HIP kernels (through the C ABI, via geomjax_amd.rmhmc) are compared with
the float64 CPU oracle (dense metric, Cholesky and autodiff as in the
reference) on the same positions, host-fed momenta draws z and uniforms u.

Tolerances (stated per result):
  * float64 kernels: 1e-10 norm-wise relative error on every state, energy
    and proposal field; fixed-point iteration counts identical; accept
    decisions identical except where |log u - dH| < 1e-8.
  * float32 kernels (FP32 state, FP64 energy/convergence test): 1e-4
    norm-wise relative error on end states; decisions identical except where
    |log u - dH| < 1e-3.
"""
import functools

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.rmhmc import RMHMCState, build_kernel, implicit_midpoint
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def oracle_chain(D, q0, eps, L, max_iters, alpha2, zs, us):
    M = om.NealFunnel(D)
    kern = orm.build_kernel(max_iters=max_iters)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    s = orm.init(q0, M.logp)
    out = []
    for z, u in zip(zs, us):
        tr = []
        s, info = kern((z, u), s, M.logp, eps, metric, L, trace=tr)
        out.append((s, info, tr))
    return out


def tie_mask(info_o, u, tol):
    dh = info_o.proposal.weight
    return (torch.log(u) - dh).abs() < tol


@pytest.mark.parametrize("D,C,eps,L,max_iters,T,seed", [
    (10, 4, 0.05, 16, 8, 3, 0),     # Neal's funnel D=10, the CPU-runnable configuration
    (100, 8, 0.03, 32, 8, 1, 1),    # D=100 shape, a few chains
    (2, 3, 0.01, 10, 100, 2, 5),    # tiny 2-D funnel
])
def test_step_f64_matches_oracle(D, C, eps, L, max_iters, T, seed):
    alpha2 = 1.0
    print(f"case D={D} C={C} eps={eps} L={L} max_iters={max_iters} T={T}", flush=True)
    q0 = seeded_normal(seed, C, D)
    zs = [seeded_normal(1000 + seed * 10 + t, C, D) for t in range(T)]
    us = [seeded_uniform(2000 + seed * 10 + t, C) for t in range(T)]
    ref = oracle_chain(D, q0, eps, L, max_iters, alpha2, zs, us)

    model = gj.NealFunnel(D)
    metric = gj.MongeMetric(alpha2)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=max_iters), return_fixed_point_info=True)
    s = gj.rmhmc.init(q0.to(DEV), model)
    ro = orm.init(q0, om.NealFunnel(D).logp)
    assert rel_err(s.logdensity, ro.logdensity) < 1e-12
    assert rel_err(s.logdensity_grad, ro.logdensity_grad) < 1e-12
    for t in range(T):
        s, info, fp = kern(gr.HostNoise(zs[t].to(DEV), us[t].to(DEV)), s, model, eps, metric, L)
        torch.cuda.synchronize()
        so, io, tr = ref[t]
        ties = tie_mask(io, us[t], 1e-8)
        assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])
        print(f"  t={t}: accepted {int(io.is_accepted.sum())}/{C}, ties {int(ties.sum())}", flush=True)
        assert rel_err(info.momentum, io.momentum) < 1e-10
        assert rel_err(info.acceptance_rate, io.acceptance_rate) < 1e-10
        assert rel_err(info.energy, io.energy) < 1e-10
        ps, pso = info.proposal.state, io.proposal.state
        for name in ("position", "momentum", "velocity", "logdensity", "logdensity_grad"):
            e = rel_err(getattr(ps, name), getattr(pso, name))
            assert e < 1e-10, (name, e)
        # dH = H0 - H1 cancels: its error is measured on the scale of the energies
        hscale = max(io.energy.abs().max().item(), 1.0)
        dw = (info.proposal.weight.cpu() - io.proposal.weight).abs().max().item()
        assert dw < 1e-10 * hscale, f"weight differs by {dw:.3e} (energy scale {hscale:.3e})"
        iters_o = torch.stack([x.iters for x in tr], dim=1)
        assert torch.equal(fp.iters.cpu().long(), iters_o)
        succ_o = torch.stack([x.success for x in tr], dim=1)
        assert torch.equal(fp.success.cpu(), succ_o)
        # the norm is a difference of iterates: its rounding error scales with the state's magnitude
        norm_o = torch.stack([x.norm for x in tr], dim=1)
        scale = max(pso.position.abs().max().item(), pso.momentum.abs().max().item(), 1.0)
        dn = (fp.norm.cpu() - norm_o).abs().max().item()
        assert dn < 1e-13 * scale * 100, f"fixed-point norm differs by {dn:.3e} (state scale {scale:.3e})"
        keep = ~ties
        assert rel_err(s.position.cpu()[keep], so.position[keep]) < 1e-10
        assert rel_err(s.logdensity.cpu()[keep], so.logdensity[keep]) < 1e-10
        assert rel_err(s.logdensity_grad.cpu()[keep], so.logdensity_grad[keep]) < 1e-10
        # continue both sides from exactly the same chain state
        s = RMHMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV))


def test_step_f32_error_vs_oracle():
    D, C, eps, L, max_iters, alpha2 = 100, 16, 0.03, 32, 8, 1.0
    q0 = seeded_normal(11, C, D)
    z = seeded_normal(12, C, D)
    u = seeded_uniform(13, C)
    so, io, _ = oracle_chain(D, q0, eps, L, max_iters, alpha2, [z], [u])[0]
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=max_iters))
    model, metric = gj.NealFunnel(D), gj.MongeMetric(alpha2)
    s = gj.rmhmc.init(q0.to(DEV).float(), model)
    s1, info = kern(gr.HostNoise(z.to(DEV).float(), u.to(DEV)), s, model, eps, metric, L)
    torch.cuda.synchronize()
    err_q = rel_err(info.proposal.state.position, io.proposal.state.position)
    err_p = rel_err(info.proposal.state.momentum, io.proposal.state.momentum)
    err_h = rel_err(info.energy, io.energy)
    print(f"float32 vs float64 oracle: position {err_q:.2e} momentum {err_p:.2e} energy {err_h:.2e}")
    assert err_q < 1e-4 and err_p < 1e-4 and err_h < 1e-4
    ties = tie_mask(io, u, 1e-3)
    assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-4)])
def test_sample_kernel_matches_step_kernel(dtype, tol):
    """The persistent T-transition launch equals T single-transition launches."""
    from geomjax_amd.sampling import run_chains

    D, C, eps, L, T = 100, 96, 0.03, 8, 4
    q0 = seeded_normal(21, C, D).to(DEV, dtype)
    zs = seeded_normal(22, T, C, D).to(DEV, dtype)
    us = seeded_uniform(23, T, C).to(DEV)
    model, metric = gj.NealFunnel(D), gj.MongeMetric(1.0)
    alg = gj.rmhmc(model, eps, metric, L, integrator=functools.partial(implicit_midpoint, max_iters=8))
    s = alg.init(q0)
    res = run_chains(alg, s, T, gr.HostNoise(zs, us))
    step_states = []
    for t in range(T):
        s, _ = alg.step(gr.HostNoise(zs[t], us[t]), s)
        step_states.append(s.position.clone())
    torch.cuda.synchronize()
    for t in range(T):
        assert rel_err(res.samples[t], step_states[t]) <= tol
    assert rel_err(res.state.position, s.position) <= tol
    assert rel_err(res.state.logdensity, s.logdensity) <= tol
    assert rel_err(res.state.logdensity_grad, s.logdensity_grad) <= tol


def test_device_prng_statistics():
    """Device-PRNG momenta: with alpha2 = 0 (G = I) the drawn momentum is z ~ N(0, I)."""
    D, C = 10, 4096
    model, metric = gj.NealFunnel(D), gj.MongeMetric(0.0)
    kern = build_kernel(functools.partial(implicit_midpoint, max_iters=8))
    s = gj.rmhmc.init(torch.zeros(C, D, dtype=torch.float64, device=DEV), model)
    _, info = kern(gr.key(7), s, model, 0.05, metric, 1)
    p = info.momentum.cpu()
    assert abs(p.mean().item()) < 0.02
    assert abs(p.var().item() - 1.0) < 0.03
    # coordinates uncorrelated
    cc = torch.corrcoef(p.T)
    off = cc - torch.eye(D, dtype=cc.dtype)
    assert off.abs().max().item() < 0.08
    # distinct keys give distinct draws, equal keys equal draws
    _, info2 = kern(gr.key(7), s, model, 0.05, metric, 1)
    _, info3 = kern(gr.key(8), s, model, 0.05, metric, 1)
    assert torch.equal(info.momentum, info2.momentum)
    assert not torch.equal(info.momentum, info3.momentum)
