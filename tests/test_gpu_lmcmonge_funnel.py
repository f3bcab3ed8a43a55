"""GPU parity: LMC with the Monge metric (lmcmonge) on Neal's funnel.

HIP kernels through the C ABI vs the float64 oracle (dense Cholesky of the
inverse metric, autodiff Hessian-vector products, the reference's
half_step_omega) on the same q0, z and u.

Tolerances: float64 kernels 1e-10 norm-wise relative on every state, info and
proposal field; dH within 1e-10 of the energy scale; accept decisions equal
except |log u - dH| < 1e-8.  float32 kernels: 1e-4.
"""
import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.lmcmonge import LMCState, build_kernel
from oracle import lmcmonge as olm
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def oracle_chain(D, q0, eps, L, alpha2, inv_mass, zs, us, faithful):
    M = om.NealFunnel(D)
    kern = olm.build_kernel(reference_faithful=faithful)
    s = olm.init(q0, M.logp)
    out = []
    for z, u in zip(zs, us):
        s, info = kern((z, u), s, M.logp, eps, inv_mass, L, alpha2)
        out.append((s, info))
    return out


CASES = [
    # D, C, eps, L, alpha2, T, seed, unit_mass, faithful
    (10, 4, 0.05, 16, 1.0, 3, 0, True, True),
    (100, 8, 0.03, 32, 1.0, 1, 1, True, True),
    (20, 6, 0.05, 8, 0.5, 2, 2, False, True),
    (20, 6, 0.05, 8, 0.5, 2, 3, False, False),
    (2, 3, 0.01, 10, 0.0, 2, 4, True, True),
]


@pytest.mark.parametrize("D,C,eps,L,alpha2,T,seed,unit_mass,faithful", CASES)
def test_step_f64_matches_oracle(D, C, eps, L, alpha2, T, seed, unit_mass, faithful):
    print(f"case D={D} C={C} eps={eps} L={L} alpha2={alpha2} T={T} unit_mass={unit_mass} faithful={faithful}",
          flush=True)
    q0 = seeded_normal(seed, C, D)
    zs = [seeded_normal(1000 + seed * 10 + t, C, D) for t in range(T)]
    us = [seeded_uniform(2000 + seed * 10 + t, C) for t in range(T)]
    inv_mass = torch.ones(D, dtype=torch.float64) if unit_mass else 0.5 + 1.5 * seeded_uniform(77, D)
    ref = oracle_chain(D, q0, eps, L, alpha2, inv_mass, zs, us, faithful)
    model = gj.NealFunnel(D)
    kern = build_kernel(reference_faithful=faithful)
    s = gj.lmcmonge.init(q0.to(DEV), model)
    for t in range(T):
        s, info = kern(gr.HostNoise(zs[t].to(DEV), us[t].to(DEV)), s, model, eps, inv_mass.to(DEV), L, alpha2)
        torch.cuda.synchronize()
        so, io = ref[t]
        ties = (torch.log(us[t]) - io.proposal.weight).abs() < 1e-8
        print(f"  t={t}: accepted {int(io.is_accepted.sum())}/{C}, ties {int(ties.sum())}", flush=True)
        assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])
        assert rel_err(info.velocity, io.velocity) < 1e-10
        assert rel_err(info.acceptance_rate, io.acceptance_rate) < 1e-10
        hscale = max(io.energy.abs().max().item(), 1.0)
        assert (info.energy.cpu() - io.energy).abs().max().item() < 1e-10 * hscale
        dw = (info.proposal.weight.cpu() - io.proposal.weight).abs().max().item()
        assert dw < 1e-10 * hscale, f"weight differs by {dw:.3e}"
        ps, pso = info.proposal.state, io.proposal.state
        # momentum = M v + alpha2 L^(faithful) (v.g) g: v.g cancels (v ~ N(0, G^-1) is nearly orthogonal to
        # g), so its error is measured on the scale of the products it sums
        g = pso.logdensity_grad_norm * torch.sqrt(pso.determinant_metric)[:, None]
        vv = -pso.velocity
        lf = pso.determinant_metric if faithful else torch.ones_like(pso.determinant_metric)
        mscale = (vv / inv_mass).abs().amax(1) + alpha2 * lf * (vv * g).abs().sum(1) * g.abs().amax(1)
        dmom = (ps.momentum.cpu() - pso.momentum).abs().amax(1)
        assert torch.all(dmom <= 1e-10 * mscale), (dmom / mscale).max().item()
        for name in ("position", "velocity", "logdensity", "logdensity_grad_norm", "dl_ig", "Hdl_ig",
                     "ig_Hdl_ig", "logdensity_hvp_velocity_norm", "determinant_metric"):
            e = rel_err(getattr(ps, name), getattr(pso, name))
            assert e < 1e-10, (name, e)
        jscale = max(pso.volume_adjustment.abs().max().item(), 1.0)
        assert (ps.volume_adjustment.cpu() - pso.volume_adjustment).abs().max().item() < 1e-10 * jscale
        keep = ~ties
        assert rel_err(s.position.cpu()[keep], so.position[keep]) < 1e-10
        assert rel_err(s.logdensity.cpu()[keep], so.logdensity[keep]) < 1e-10
        assert rel_err(s.logdensity_grad.cpu()[keep], so.logdensity_grad[keep]) < 1e-10
        s = LMCState(so.position.to(DEV), so.logdensity.to(DEV), so.logdensity_grad.to(DEV),
                     so.volume_adjustment.to(DEV))


def test_step_f32_error_vs_oracle():
    D, C, eps, L, alpha2 = 100, 16, 0.03, 32, 1.0
    q0, z, u = seeded_normal(31, C, D), seeded_normal(32, C, D), seeded_uniform(33, C)
    inv_mass = torch.ones(D, dtype=torch.float64)
    so, io = oracle_chain(D, q0, eps, L, alpha2, inv_mass, [z], [u], True)[0]
    model = gj.NealFunnel(D)
    s = gj.lmcmonge.init(q0.to(DEV).float(), model)
    s1, info = build_kernel()(gr.HostNoise(z.to(DEV).float(), u.to(DEV)), s, model, eps,
                              inv_mass.to(DEV).float(), L, alpha2)
    torch.cuda.synchronize()
    eq = rel_err(info.proposal.state.position, io.proposal.state.position)
    ev = rel_err(info.proposal.state.velocity, io.proposal.state.velocity)
    eh = rel_err(info.energy, io.energy)
    print(f"float32 vs float64 oracle: position {eq:.2e} velocity {ev:.2e} energy {eh:.2e}")
    assert eq < 1e-4 and ev < 1e-4 and eh < 1e-4
    ties = (torch.log(u) - io.proposal.weight).abs() < 1e-3
    assert torch.equal(info.is_accepted.cpu()[~ties], io.is_accepted[~ties])


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-4)])
def test_sample_kernel_matches_step_kernel(dtype, tol):
    from geomjax_amd.sampling import run_chains

    D, C, eps, L, T = 100, 96, 0.03, 8, 4
    q0 = seeded_normal(41, C, D).to(DEV, dtype)
    zs = seeded_normal(42, T, C, D).to(DEV, dtype)
    us = seeded_uniform(43, T, C).to(DEV)
    alg = gj.lmcmonge(gj.NealFunnel(D), eps, torch.ones(D, dtype=dtype, device=DEV), L, alpha2=1.0)
    s = alg.init(q0)
    res = run_chains(alg, s, T, gr.HostNoise(zs, us))
    for t in range(T):
        s, _ = alg.step(gr.HostNoise(zs[t], us[t]), s)
        assert rel_err(res.samples[t], s.position) <= tol
    torch.cuda.synchronize()
    assert rel_err(res.state.position, s.position) <= tol
    assert rel_err(res.state.logdensity, s.logdensity) <= tol
    assert rel_err(res.state.volume_adjustment, s.volume_adjustment) <= max(tol, 1e-12)
