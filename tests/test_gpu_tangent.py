"""GPU: derivatives of ESJD x acceptance w.r.t. (log step_size, log alpha2).

The tangent kernel is checked (float64) against
  * torch.autograd through the float64 oracle (dense metric, autodiff),
    unrolled over the same executed fixed-point iterations: 1e-8 relative;
  * central finite differences (h = 1e-5) of the primal HIP step kernel with
    the iteration count fixed (convergence_tol = 0, so every step runs
    max_iters iterations and the objective is smooth in the parameters):
    1e-6 relative.
"""
import functools
import math

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.adjoint import ESJDAcceptance, esjd_acceptance
from geomjax_amd.rmhmc import build_kernel, implicit_midpoint
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def oracle_objective_grad(D, q0, z, eps, alpha2, L, max_iters, conv_tol=1e-6):
    le = torch.tensor(math.log(eps), dtype=torch.float64, requires_grad=True)
    la = torch.tensor(math.log(alpha2), dtype=torch.float64, requires_grad=True)
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, torch.exp(la), D)
    k = orm.build_kernel(max_iters=max_iters, convergence_tol=conv_tol)
    u = torch.zeros(q0.shape[0], dtype=torch.float64)
    _, info = k((z, u), orm.init(q0, M.logp), M.logp, torch.exp(le), metric, L)
    jump = ((info.proposal.state.position - q0) ** 2).sum(-1)
    obj = info.acceptance_rate * jump
    g = [torch.autograd.grad(o, (le, la), retain_graph=True) for o in obj]
    return obj.detach(), torch.stack([torch.stack(x) for x in g])


@pytest.mark.parametrize("D,C,eps,alpha2,L,max_iters", [(10, 4, 0.05, 1.0, 8, 8), (30, 3, 0.03, 0.6, 6, 8)])
def test_tangent_matches_autograd_through_oracle(D, C, eps, alpha2, L, max_iters):
    q0, z = seeded_normal(90 + D, C, D) * 0.8, seeded_normal(91 + D, C, D)
    obj_o, g_o = oracle_objective_grad(D, q0, z, eps, alpha2, L, max_iters)
    model = gj.NealFunnel(D)
    s = gj.rmhmc.init(q0.to(DEV), model)
    r = esjd_acceptance(gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV)), s, model, eps, gj.MongeMetric(alpha2),
                        L, integrator=functools.partial(implicit_midpoint, max_iters=max_iters))
    torch.cuda.synchronize()
    print("objective", r.objective.cpu().tolist(), "grad", r.d_objective.cpu().tolist(), "oracle", g_o.tolist())
    assert rel_err(r.objective, obj_o) < 1e-10
    assert rel_err(r.d_objective, g_o) < 1e-8


def test_tangent_matches_central_differences():
    D, C, eps, alpha2, L, max_iters, h = 20, 6, 0.04, 0.8, 8, 6, 1e-5
    q0, z = seeded_normal(95, C, D) * 0.8, seeded_normal(96, C, D)
    model = gj.NealFunnel(D)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters, convergence_tol=0.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    noise = gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV, dtype=torch.float64))
    r = esjd_acceptance(noise, s, model, eps, gj.MongeMetric(alpha2), L, integrator=integ)
    kern = build_kernel(integ)

    def obj(le, la):
        _, info = kern(noise, s, model, math.exp(le), gj.MongeMetric(math.exp(la)), L)
        jump = ((info.proposal.state.position - s.position) ** 2).sum(-1)
        return (info.acceptance_rate * jump).double()

    le, la = math.log(eps), math.log(alpha2)
    fd = torch.stack([(obj(le + h, la) - obj(le - h, la)) / (2 * h), (obj(le, la + h) - obj(le, la - h)) / (2 * h)],
                     dim=1)
    torch.cuda.synchronize()
    print("tangent", r.d_objective.cpu().tolist(), "fd", fd.cpu().tolist())
    assert rel_err(r.objective, obj(le, la)) < 1e-12
    assert rel_err(r.d_objective, fd) < 1e-6


def test_autograd_function():
    D, C = 10, 64
    model = gj.NealFunnel(D)
    s = gj.rmhmc.init(seeded_normal(97, C, D).to(DEV), model)
    key = gr.key(3)
    le = torch.tensor(math.log(0.05), dtype=torch.float64, requires_grad=True)
    la = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    val = ESJDAcceptance.apply(le, la, s, key, model, 8)
    val.backward()
    r = esjd_acceptance(key, s, model, 0.05, gj.MongeMetric(1.0), 8)
    mean = r.d_objective.mean(0).cpu()
    assert abs(le.grad.item() - mean[0].item()) <= 1e-12 * max(1.0, abs(mean[0].item()))
    assert abs(la.grad.item() - mean[1].item()) <= 1e-12 * max(1.0, abs(mean[1].item()))


def oracle_logistic_objective_grad(X, y, q0, z, eps, L, max_iters):
    le = torch.tensor(math.log(eps), dtype=torch.float64, requires_grad=True)
    M = om.LogisticRegression(X, y, 100.0)
    k = orm.build_kernel(max_iters=max_iters)
    u = torch.zeros(q0.shape[0], dtype=torch.float64)
    _, info = k((z, u), orm.init(q0, M.logp), M.logp, torch.exp(le), M.fisher_metric_fn, L)
    jump = ((info.proposal.state.position - q0) ** 2).sum(-1)
    obj = info.acceptance_rate * jump
    g = [torch.autograd.grad(o, (le,), retain_graph=True)[0] for o in obj]
    return obj.detach(), torch.stack(g)


@pytest.mark.parametrize("N,D,C,eps,L,max_iters,seed", [
    (256, 8, 3, 0.1, 6, 6, 0),
    (300, 13, 2, 0.15, 5, 6, 1),    # ragged N and D
    (1024, 64, 2, 0.1, 4, 6, 2),    # at the tangent kernel's capacity
])
def test_logistic_tangent_matches_autograd_through_oracle(N, D, C, eps, L, max_iters, seed):
    from geomjax_amd.synthetic import make_logistic

    print(f"case N={N} D={D} C={C} eps={eps} L={L}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    q0, z = seeded_normal(120 + seed, C, D) * 0.5, seeded_normal(121 + seed, C, D)
    obj_o, g_o = oracle_logistic_objective_grad(X, y, q0, z, eps, L, max_iters)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    r = esjd_acceptance(gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV)), s, model, eps, gj.FisherMetric(), L,
                        integrator=functools.partial(implicit_midpoint, max_iters=max_iters))
    torch.cuda.synchronize()
    print("  objective", r.objective.cpu().tolist(), "d/dlog eps", r.d_objective[:, 0].cpu().tolist(),
          "oracle", g_o.tolist(), flush=True)
    assert rel_err(r.objective, obj_o) < 1e-10
    assert rel_err(r.d_objective[:, 0], g_o) < 1e-8
    assert float(r.d_objective[:, 1].abs().max()) == 0.0
    assert rel_err(r.d_objective[:, 0], r.d_acceptance[:, 0] * r.sq_jump + r.acceptance * r.d_sq_jump[:, 0]) < 1e-12


def test_logistic_tangent_matches_central_differences():
    from geomjax_amd.synthetic import make_logistic

    N, D, C, eps, L, max_iters, h = 512, 24, 4, 0.12, 6, 5, 1e-5
    X, y, _ = make_logistic(N, D, 5)
    q0, z = seeded_normal(130, C, D) * 0.5, seeded_normal(131, C, D)
    model = gj.LogisticRegression(X, y, 100.0)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters, convergence_tol=0.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    noise = gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV, dtype=torch.float64))
    r = esjd_acceptance(noise, s, model, eps, gj.FisherMetric(), L, integrator=integ)
    kern = build_kernel(integ)

    def obj(le):
        _, info = kern(noise, s, model, math.exp(le), gj.FisherMetric(), L)
        jump = ((info.proposal.state.position - s.position) ** 2).sum(-1)
        return (info.acceptance_rate * jump).double()

    le = math.log(eps)
    fd = (obj(le + h) - obj(le - h)) / (2 * h)
    torch.cuda.synchronize()
    print("tangent", r.d_objective[:, 0].cpu().tolist(), "fd", fd.cpu().tolist())
    assert rel_err(r.objective, obj(le)) < 1e-12
    assert rel_err(r.d_objective[:, 0], fd) < 1e-6


def test_logistic_autograd_function_and_adapt():
    from geomjax_amd.adjoint import adapt
    from geomjax_amd.synthetic import make_logistic

    N, D, C = 512, 16, 32
    X, y, _ = make_logistic(N, D, 6)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.rmhmc.init(seeded_normal(140, C, D).to(DEV) * 0.3, model)
    key = gr.key(4)
    le = torch.tensor(math.log(0.1), dtype=torch.float64, requires_grad=True)
    la = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    val = ESJDAcceptance.apply(le, la, s, key, model, 6)
    val.backward()
    r = esjd_acceptance(key, s, model, 0.1, gj.FisherMetric(), 6)
    mean = r.d_objective.mean(0).cpu()
    assert abs(le.grad.item() - mean[0].item()) <= 1e-12 * max(1.0, abs(mean[0].item()))
    assert la.grad.item() == 0.0
    eps, a2, s2, hist = adapt(key, s, model, 6, step_size=0.02, num_steps=15, learning_rate=0.1)
    print("adapted step size", eps, "history", [round(h[2], 4) for h in hist])
    assert a2 == 1.0 and math.isfinite(eps) and eps > 0.02
    assert torch.isfinite(s2.position).all()


# ---------------------------------------------------------------------------
# The adjoint configuration's own sizes: 32 integration steps on the funnel
# with D = 100 (eps 0.03, alpha2 1) and on logistic regression 2048 x 64
# (eps 0.1).  Derivative error can grow with trajectory length, so the bars
# (1e-8 vs autograd through the oracle, 1e-6 vs central differences) are
# checked at the full length, a few chains each.
# ---------------------------------------------------------------------------

def test_tangent_funnel_d100_32_steps_matches_autograd():
    D, C, eps, alpha2, L, max_iters = 100, 3, 0.03, 1.0, 32, 8
    print(f"case funnel D={D} C={C} L={L} vs autograd", flush=True)
    q0, z = seeded_normal(200, C, D) * 0.8, seeded_normal(201, C, D)
    obj_o, g_o = oracle_objective_grad(D, q0, z, eps, alpha2, L, max_iters)
    model = gj.NealFunnel(D)
    s = gj.rmhmc.init(q0.to(DEV), model)
    r = esjd_acceptance(gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV)), s, model, eps, gj.MongeMetric(alpha2),
                        L, integrator=functools.partial(implicit_midpoint, max_iters=max_iters))
    torch.cuda.synchronize()
    print("  objective", r.objective.cpu().tolist(), "grad", r.d_objective.cpu().tolist(), "oracle", g_o.tolist(),
          flush=True)
    assert rel_err(r.objective, obj_o) < 1e-10
    assert rel_err(r.d_objective, g_o) < 1e-8


def test_tangent_funnel_d100_32_steps_matches_central_differences():
    """Central differences at h = 1e-5 with one Richardson step (h and 2h):
    32 steps of eight fixed-point iterations each make the objective's third
    derivative large enough on some chains that the plain central difference's
    O(h^2) truncation error alone exceeds 1e-6 relative; the extrapolated
    difference is O(h^4)."""
    D, C, eps, alpha2, L, max_iters, h = 100, 6, 0.03, 1.0, 32, 8, 1e-5
    print(f"case funnel D={D} C={C} L={L} vs central differences", flush=True)
    q0, z = seeded_normal(202, C, D) * 0.8, seeded_normal(203, C, D)
    model = gj.NealFunnel(D)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters, convergence_tol=0.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    noise = gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV, dtype=torch.float64))
    r = esjd_acceptance(noise, s, model, eps, gj.MongeMetric(alpha2), L, integrator=integ)
    kern = build_kernel(integ)

    def obj(le, la):
        _, info = kern(noise, s, model, math.exp(le), gj.MongeMetric(math.exp(la)), L)
        jump = ((info.proposal.state.position - s.position) ** 2).sum(-1)
        return (info.acceptance_rate * jump).double()

    le, la = math.log(eps), math.log(alpha2)

    def central(hh):
        return torch.stack([(obj(le + hh, la) - obj(le - hh, la)) / (2 * hh),
                            (obj(le, la + hh) - obj(le, la - hh)) / (2 * hh)], dim=1)

    fd1, fd2 = central(h), central(2 * h)
    fd = (4.0 * fd1 - fd2) / 3.0
    torch.cuda.synchronize()
    print("  tangent", r.d_objective.cpu().tolist(), "\n  fd(h)", fd1.cpu().tolist(), "\n  fd(2h)",
          fd2.cpu().tolist(), "\n  richardson", fd.cpu().tolist(), flush=True)
    print(f"  relative error vs fd(h) {rel_err(r.d_objective, fd1):.2e}, vs Richardson {rel_err(r.d_objective, fd):.2e}",
          flush=True)
    assert rel_err(r.objective, obj(le, la)) < 1e-12
    assert rel_err(r.d_objective, fd) < 1e-6


def test_logistic_tangent_2048x64_32_steps_matches_autograd():
    from geomjax_amd.synthetic import make_logistic

    N, D, C, eps, L, max_iters = 2048, 64, 2, 0.1, 32, 6
    print(f"case logistic {N}x{D} C={C} L={L} vs autograd", flush=True)
    X, y, _ = make_logistic(N, D, 2)
    q0, z = seeded_normal(210, C, D) * 0.3, seeded_normal(211, C, D)
    obj_o, g_o = oracle_logistic_objective_grad(X, y, q0, z, eps, L, max_iters)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    r = esjd_acceptance(gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV)), s, model, eps, gj.FisherMetric(), L,
                        integrator=functools.partial(implicit_midpoint, max_iters=max_iters))
    torch.cuda.synchronize()
    print("  objective", r.objective.cpu().tolist(), "d/dlog eps", r.d_objective[:, 0].cpu().tolist(),
          "oracle", g_o.tolist(), flush=True)
    assert rel_err(r.objective, obj_o) < 1e-10
    assert rel_err(r.d_objective[:, 0], g_o) < 1e-8


def test_logistic_tangent_2048x64_32_steps_matches_central_differences():
    from geomjax_amd.synthetic import make_logistic

    N, D, C, eps, L, max_iters, h = 2048, 64, 4, 0.1, 32, 6, 1e-5
    print(f"case logistic {N}x{D} C={C} L={L} vs central differences", flush=True)
    X, y, _ = make_logistic(N, D, 2)
    q0, z = seeded_normal(212, C, D) * 0.3, seeded_normal(213, C, D)
    model = gj.LogisticRegression(X, y, 100.0)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters, convergence_tol=0.0)
    s = gj.rmhmc.init(q0.to(DEV), model)
    noise = gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV, dtype=torch.float64))
    r = esjd_acceptance(noise, s, model, eps, gj.FisherMetric(), L, integrator=integ)
    kern = build_kernel(integ)

    def obj(le):
        _, info = kern(noise, s, model, math.exp(le), gj.FisherMetric(), L)
        jump = ((info.proposal.state.position - s.position) ** 2).sum(-1)
        return (info.acceptance_rate * jump).double()

    le = math.log(eps)
    fd = (obj(le + h) - obj(le - h)) / (2 * h)
    torch.cuda.synchronize()
    print("  tangent", r.d_objective[:, 0].cpu().tolist(), "fd", fd.cpu().tolist(), flush=True)
    assert rel_err(r.objective, obj(le)) < 1e-12
    assert rel_err(r.d_objective[:, 0], fd) < 1e-6


@pytest.mark.parametrize("D,C,eps,alpha2,L,max_iters", [(10, 16, 0.05, 1.0, 8, 8), (100, 16, 0.03, 1.0, 32, 8)])
def test_funnel_fast_path_float32_matches_float64(D, C, eps, alpha2, L, max_iters):
    """The float32 fast path (float32 state, trajectory and tangents, float64
    energies) against the float64 check path on the same positions and normals
    (the check path is pinned above against autograd through the oracle and
    central differences): objective 1e-4 and derivatives 5e-3, relative over
    the chains (observed at D = 100 with 32 steps: 2.7e-5 and up to 9.5e-4 --
    float32 rounding carried through 32 x 8 fixed-point iterates)."""
    print(f"case funnel D={D} C={C} L={L} float32 fast path", flush=True)
    model = gj.NealFunnel(D)
    q0, z = seeded_normal(300, C, D) * 0.8, seeded_normal(301, C, D)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters)
    s64 = gj.rmhmc.init(q0.to(DEV), model)
    s32 = gj.rmhmc.init(q0.to(DEV).float(), model)
    r64 = esjd_acceptance(gr.HostNoise(z.to(DEV), torch.zeros(C, device=DEV)), s64, model, eps,
                          gj.MongeMetric(alpha2), L, integrator=integ)
    r32 = esjd_acceptance(gr.HostNoise(z.to(DEV).float(), torch.zeros(C, device=DEV)), s32, model, eps,
                          gj.MongeMetric(alpha2), L, integrator=integ)
    torch.cuda.synchronize()
    e_obj = rel_err(r32.objective, r64.objective)
    e_d = rel_err(r32.d_objective, r64.d_objective)
    e_acc = rel_err(r32.d_acceptance, r64.d_acceptance)
    print(f"  objective {e_obj:.2e}  d objective {e_d:.2e}  d acceptance {e_acc:.2e}", flush=True)
    assert r32.objective.dtype == torch.float64
    assert e_obj < 1e-4
    assert e_d < 5e-3
    assert e_acc < 5e-3


@pytest.mark.parametrize("N,D,C,eps,L", [(512, 8, 32, 0.1, 8), (1024, 40, 32, 0.1, 16), (2048, 64, 32, 0.1, 32)])
def test_logistic_fast_path_float32_matches_float64(N, D, C, eps, L):
    """The logistic float32 fast path (the tuned kernel's float32 state and
    split-bf16 MFMA passes with tangents, float64 energies) against the float64
    check path (pinned above against autograd through the oracle and central
    differences) on the same float32-representable positions and normals, up
    to the tangent configuration (2048 x 64, 32 steps): objective 2e-3 and its
    derivative, the acceptance's derivative and the squared jump's derivative
    1e-2, relative over the chains (observed at 2048 x 64: 7.5e-4 and up to
    2.2e-3; the float32 fixed-point norms stop some steps one iteration earlier
    or later than float64, so the two differentiate slightly different
    executed computations)."""
    from geomjax_amd.synthetic import make_logistic

    print(f"case logistic N={N} D={D} C={C} L={L} float32 fast path", flush=True)
    X, y, _ = make_logistic(N, D, 70 + D)
    model = gj.LogisticRegression(X, y, 100.0)
    q0 = (0.3 * seeded_normal(310 + D, C, D)).float().double()
    z = seeded_normal(311 + D, C, D).float().double()
    integ = functools.partial(implicit_midpoint, max_iters=6)
    u0 = torch.zeros(C, device=DEV)
    r64 = esjd_acceptance(gr.HostNoise(z.to(DEV), u0), gj.rmhmc.init(q0.to(DEV), model), model, eps,
                          gj.FisherMetric(), L, integrator=integ)
    r32 = esjd_acceptance(gr.HostNoise(z.to(DEV).float(), u0), gj.rmhmc.init(q0.to(DEV).float(), model), model, eps,
                          gj.FisherMetric(), L, integrator=integ)
    torch.cuda.synchronize()
    errs = {"objective": rel_err(r32.objective, r64.objective),
            "d objective": rel_err(r32.d_objective[:, 0], r64.d_objective[:, 0]),
            "d acceptance": rel_err(r32.d_acceptance[:, 0], r64.d_acceptance[:, 0]),
            "d sq jump": rel_err(r32.d_sq_jump[:, 0], r64.d_sq_jump[:, 0])}
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
          f"; acceptance {float(r64.acceptance.mean()):.4f}; fevals {r32.num_fevals.tolist()[:4]} vs "
          f"{r64.num_fevals.tolist()[:4]}", flush=True)
    assert r32.objective.dtype == torch.float64
    assert torch.all(r32.d_objective[:, 1] == 0)
    assert errs["objective"] < 2e-3
    for k in ("d objective", "d acceptance", "d sq jump"):
        assert errs[k] < 1e-2, k


@pytest.mark.parametrize("N,D,C,eps,L,max_iters", [(300, 100, 3, 0.1, 4, 5), (512, 128, 2, 0.08, 3, 4)])
def test_logistic_wide_tangent_float32_matches_autograd_through_oracle(N, D, C, eps, L, max_iters):
    """The tangent for 65..128 features (float32 state and tangents on fp32
    MFMA, float64 energies and objective; no float64 tangent holds G^-1 and its
    tangent at this width) against autograd through the float64 oracle from the
    same float32-representable positions and normals, with convergence_tol = 0
    on both sides so that every step runs the same number of fixed-point
    iterations: objective 1e-4 and derivative 1e-3 relative over the chains
    (observed 5.5e-6 / 6.8e-6 at D = 100 and 4.6e-7 / 4.8e-7 at D = 128)."""
    from geomjax_amd.synthetic import make_logistic

    print(f"case wide tangent N={N} D={D} C={C} eps={eps} L={L}", flush=True)
    X, y, _ = make_logistic(N, D, 90 + D)
    q0 = (0.3 * seeded_normal(320 + D, C, D)).float().double()
    z = seeded_normal(321 + D, C, D).float().double()
    le = torch.tensor(math.log(eps), dtype=torch.float64, requires_grad=True)
    Mo = om.LogisticRegression(X, y, 100.0)
    k = orm.build_kernel(max_iters=max_iters, convergence_tol=0.0)
    _, info = k((z, torch.zeros(C, dtype=torch.float64)), orm.init(q0, Mo.logp), Mo.logp, torch.exp(le),
                Mo.fisher_metric_fn, L)
    obj_o = info.acceptance_rate * ((info.proposal.state.position - q0) ** 2).sum(-1)
    g_o = torch.stack([torch.autograd.grad(o, (le,), retain_graph=True)[0] for o in obj_o])
    model = gj.LogisticRegression(X, y, 100.0)
    r = esjd_acceptance(gr.HostNoise(z.to(DEV).float(), torch.zeros(C, device=DEV)),
                        gj.rmhmc.init(q0.to(DEV).float(), model), model, eps, gj.FisherMetric(), L,
                        integrator=functools.partial(implicit_midpoint, max_iters=max_iters, convergence_tol=0.0))
    torch.cuda.synchronize()
    e_obj = rel_err(r.objective, obj_o.detach())
    e_d = rel_err(r.d_objective[:, 0], g_o)
    print(f"  objective {r.objective.cpu().tolist()} vs {obj_o.detach().tolist()} ({e_obj:.1e}); "
          f"d/dlog eps {r.d_objective[:, 0].cpu().tolist()} vs {g_o.tolist()} ({e_d:.1e}); "
          f"fevals {r.num_fevals.tolist()}", flush=True)
    assert float(r.d_objective[:, 1].abs().max()) == 0.0
    assert e_obj < 1e-4
    assert e_d < 1e-3


def test_adapt_a_batch_with_wide_posteriors_in_float32():
    """Step-size tuning by gradient on a batch of posteriors of 24 and 100
    features with float32 states (the one-wave fast path serves the narrow
    posterior, the 65..128 float32 tangent the wide one): the derivative is
    finite for every chain and Adam raises a too-small step size."""
    from geomjax_amd.adjoint import adapt
    from geomjax_amd.synthetic import make_logistic

    models = [gj.LogisticRegression(*make_logistic(400, d, 60 + d)[:2], 100.0) for d in (24, 100)]
    cpp = 16
    batch = gj.LogisticBatch(models, chains_per_posterior=cpp)
    q = torch.zeros(2 * cpp, batch.dim)
    q[:cpp, :24] = 0.3 * seeded_normal(150, cpp, 24)
    q[cpp:, :100] = 0.3 * seeded_normal(151, cpp, 100)
    s = gj.rmhmc.init(q.to(DEV).float(), batch)
    r = esjd_acceptance(gr.key(5), s, batch, 0.02, gj.FisherMetric(), 6)
    torch.cuda.synchronize()
    print("d objective per posterior", r.d_objective[:cpp, 0].mean().item(), r.d_objective[cpp:, 0].mean().item(),
          flush=True)
    assert torch.isfinite(r.d_objective).all() and torch.isfinite(r.objective).all()
    assert torch.all(r.num_fevals > 0)
    eps, a2, s2, hist = adapt(gr.key(6), s, batch, 6, step_size=0.02, num_steps=12, learning_rate=0.1)
    print("adapted step size", eps, "history", [round(h[2], 4) for h in hist], flush=True)
    assert math.isfinite(eps) and eps > 0.02
    assert s2.position.dtype == torch.float32 and torch.isfinite(s2.position).all()
