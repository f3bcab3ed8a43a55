"""Pins of the oracle's RMHMC restatement (CPU, float64).

None of these retypes the oracle's formulas: each compares the oracle against
an independent fact (the reference's own cross-sampler invariant, a closed
form, a textbook map, finite differences, a symplectic invariant).
"""
import math

import pytest
import torch
from torch.func import grad, vmap

from oracle import mcmc as omc
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform

F64 = torch.float64


def test_identity_metric_rmhmc_matches_leapfrog_hmc():
    """tests/test_samplers.py:21-32,56: rmhmc(G=I) == hmc to rtol 1e-4 (h=1e-2, L=10, zeros(2))."""
    D = 2
    M = om.NealFunnel(D)
    z, u = seeded_normal(0, 3, D), seeded_uniform(1, 3)
    q0 = torch.zeros(3, D, dtype=F64)
    init_fn, step_fn = orm.rmhmc(M.logp, 1e-2, om.identity_metric_fn(D), 10)
    s1, _ = step_fn((z, u), init_fn(q0))
    h1, _ = omc.hmc_kernel((z, u), omc.hmc_init(q0, M.logp), M.logp, 1e-2, torch.ones(D, dtype=F64), 10)
    assert torch.allclose(s1.position, h1.position, rtol=1e-4)
    # and the two integrators are genuinely different schemes (not the same code)
    assert (s1.position - h1.position).abs().max() > 0


def _monge_dTdq_closed_form(logp, alpha2, q, p):
    """dT/dq for G = I + a2 g g^T from Sherman-Morrison and the determinant lemma."""
    g = grad(logp)(q)
    H = torch.func.hessian(logp)(q)
    L = 1 + alpha2 * g @ g
    u = p - alpha2 * g * (g @ p) / L
    return alpha2 * (H @ g / L - (g @ u) * (H @ u)), u


@pytest.mark.parametrize("D,alpha2", [(3, 1.0), (6, 0.37)])
def test_monge_kinetic_gradient_closed_form(D, alpha2):
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    _, ke, imvp = orm.gaussian_riemannian(metric)
    for s in range(3):
        q, p = seeded_normal(10 + s, D), seeded_normal(20 + s, D)
        dTdq, dTdp = grad(ke, argnums=(0, 1))(q, p)
        cf_q, cf_u = _monge_dTdq_closed_form(M.logp, alpha2, q, p)
        assert rel_err(dTdq, cf_q) < 1e-12
        assert rel_err(dTdp, cf_u) < 1e-12
        assert rel_err(imvp(q, p), cf_u) < 1e-12
        # kinetic energy value: 0.5 p^T G^-1 p + 0.5 log(1 + a2 |g|^2) + (D/2) log(2 pi)
        g = grad(M.logp)(q)
        L = 1 + alpha2 * g @ g
        val = 0.5 * (p @ p - alpha2 * (g @ p) ** 2 / L) + 0.5 * torch.log(L) + 0.5 * D * math.log(2 * math.pi)
        assert abs(ke(q, p) - val) < 1e-12 * abs(val)


def test_kinetic_gradient_central_differences():
    D, alpha2 = 4, 0.8
    M = om.NealFunnel(D)
    _, ke, _ = orm.gaussian_riemannian(om.monge_metric_fn(M.logp, alpha2, D))
    q, p = seeded_normal(3, D) * 0.5, seeded_normal(4, D)
    dTdq = grad(ke, argnums=0)(q, p)
    h = 1e-6
    fd = torch.stack([(ke(q + h * e, p) - ke(q - h * e, p)) / (2 * h) for e in torch.eye(D, dtype=F64)])
    assert rel_err(dTdq, fd) < 1e-7


def test_funnel_gradient_closed_form():
    D = 7
    M = om.NealFunnel(D)
    q = seeded_normal(5, D)
    g = grad(M.logp)(q)
    x, v = q[:-1], q[-1]
    gx = -x * torch.exp(-v)
    gv = -v / 9.0 - (D - 1) / 2 + 0.5 * torch.exp(-v) * (x @ x)
    assert rel_err(g, torch.cat([gx, gv[None]])) < 1e-13
    # log density: v ~ N(0, 9), x | v ~ N(0, e^v)
    lp = -v * v / 18 - 0.5 * math.log(18 * math.pi) + (-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi)).sum()
    assert abs(M.logp(q) - lp) < 1e-12


def test_momentum_draw_is_cholesky_factor():
    """p = chol(G) z with G = I + a2 g g^T: check E[p p^T] = G through the factor itself."""
    D, alpha2 = 5, 1.0
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    mg, _, _ = orm.gaussian_riemannian(metric)
    q = seeded_normal(6, D)
    G = metric(q)
    cols = torch.stack([mg(e, q) for e in torch.eye(D, dtype=F64)], dim=1)  # the factor L
    assert rel_err(cols @ cols.T, G) < 1e-13
    assert torch.all(torch.triu(cols, 1).abs() < 1e-15)


# --------------------------------------------------------------------------
# fixed-point solver semantics (geomjax/rmhmc/integrators.py:53-89)
# --------------------------------------------------------------------------

def _lin(a, b):
    def f(x):
        q, p = x
        return (a * q + b, a * p + b), q * 0
    return f


def test_fixed_point_counts_and_stop_rules():
    # x_{n+1} = 0.5 x_n + 1 from 0: x_n = 2 - 2^{1-n}; |x_{n+1} - x_n| = 2^{-n}
    x0 = (torch.zeros(1, 1, dtype=F64), torch.zeros(1, 1, dtype=F64))
    x, _, info = orm.solve_fixed_point_iteration(_lin(0.5, 1.0), x0, convergence_tol=1e-3, max_iters=100)
    # first call gives x1 = 1 (norm 1, outside the loop), loop until 2^{-n} <= 1e-3 -> n = 10 evaluations in the loop
    assert int(info.iters) == 10
    assert bool(info.success)
    assert abs(float(info.norm) - 2.0 ** -10) < 1e-15
    assert abs(float(x[0]) - (2 - 2.0 ** -10)) < 1e-15
    # max_iters cap
    _, _, info = orm.solve_fixed_point_iteration(_lin(0.5, 1.0), x0, convergence_tol=1e-3, max_iters=3)
    assert int(info.iters) == 3 and not bool(info.success)
    # divergence: x_{n+1} = 10 x_n + 1 stops once the norm reaches 1e10
    _, _, info = orm.solve_fixed_point_iteration(_lin(10.0, 1.0), x0, divergence_tol=1e10, max_iters=100)
    assert float(info.norm) >= 1e10 and not bool(info.success)
    # NaN stops immediately after the first evaluation
    _, _, info = orm.solve_fixed_point_iteration(_lin(float("nan"), 1.0),
                                                 (torch.ones(1, 1, dtype=F64), torch.ones(1, 1, dtype=F64)))
    assert int(info.iters) == 0 and not bool(info.success)


def test_fixed_point_batched_freeze():
    """Chains stop independently (batched while_loop semantics)."""
    a = torch.tensor([[0.5], [0.1]], dtype=F64)
    f = lambda x: ((a * x[0] + 1, a * x[1] + 1), x[0] * 0)
    x0 = (torch.zeros(2, 1, dtype=F64), torch.zeros(2, 1, dtype=F64))
    x, _, info = orm.solve_fixed_point_iteration(f, x0, convergence_tol=1e-3, max_iters=100)
    xa, _, ia = orm.solve_fixed_point_iteration(lambda y: ((0.5 * y[0] + 1, 0.5 * y[1] + 1), y[0] * 0),
                                                (x0[0][:1], x0[1][:1]), convergence_tol=1e-3)
    xb, _, ib = orm.solve_fixed_point_iteration(lambda y: ((0.1 * y[0] + 1, 0.1 * y[1] + 1), y[0] * 0),
                                                (x0[0][:1], x0[1][:1]), convergence_tol=1e-3)
    assert info.iters.tolist() == [int(ia.iters), int(ib.iters)]
    assert float(x[0][0]) == float(xa[0][0]) and float(x[0][1]) == float(xb[0][0])


# --------------------------------------------------------------------------
# integrator invariants
# --------------------------------------------------------------------------

def _gaussian_problem(D, seed):
    A = seeded_normal(seed, D, D)
    A = A @ A.T / D + torch.eye(D, dtype=F64)
    Gm = seeded_normal(seed + 1, D, D)
    Gm = Gm @ Gm.T / D + 0.5 * torch.eye(D, dtype=F64)
    logp = lambda q: -0.5 * q @ (A @ q)
    metric = lambda q: Gm + 0.0 * q.sum()
    return A, Gm, logp, metric


def test_implicit_midpoint_is_cayley_map_for_gaussian_constant_metric():
    D, h = 4, 0.3
    A, Gm, logp, metric = _gaussian_problem(D, 40)
    _, ke, imvp = orm.gaussian_riemannian(metric)
    step = orm.implicit_midpoint(logp, ke, imvp, convergence_tol=1e-15, max_iters=500)
    q0, p0 = seeded_normal(41, 2, D), seeded_normal(42, 2, D)
    g0 = vmap(grad(logp))(q0)
    st = orm.IntegratorState(q0, p0, p0, torch.zeros(2, dtype=F64), g0)
    out = step(st, h)
    # linear system z' = M z, M = [[0, G^-1], [-A, 0]]; midpoint => Cayley map
    Z = torch.zeros(D, D, dtype=F64)
    Mm = torch.cat([torch.cat([Z, torch.linalg.inv(Gm)], 1), torch.cat([-A, Z], 1)], 0)
    I2 = torch.eye(2 * D, dtype=F64)
    cay = torch.linalg.solve(I2 - 0.5 * h * Mm, I2 + 0.5 * h * Mm)
    z1 = (cay @ torch.cat([q0, p0], 1).T).T
    assert rel_err(torch.cat([out.position, out.momentum], 1), z1) < 1e-12


def _monge_step(D, alpha2, tol=1e-14, max_iters=400):
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    _, ke, imvp = orm.gaussian_riemannian(metric)
    return M, orm.implicit_midpoint(M.logp, ke, imvp, convergence_tol=tol, max_iters=max_iters), ke


def test_reversibility():
    D, h, L = 5, 0.05, 12
    M, step, _ = _monge_step(D, 1.0)
    q0, p0 = seeded_normal(50, 3, D) * 0.7, seeded_normal(51, 3, D)
    st = orm.IntegratorState(q0, p0, p0, torch.zeros(3, dtype=F64), vmap(grad(M.logp))(q0))
    for _ in range(L):
        st = step(st, h)
    st = orm.flip_momentum(st)
    for _ in range(L):
        st = step(st, h)
    assert rel_err(st.position, q0) < 1e-10
    assert rel_err(-st.momentum, p0) < 1e-10


def test_symplectic_unit_jacobian():
    D, h = 3, 0.1
    M, step, _ = _monge_step(D, 1.0)
    z = torch.cat([seeded_normal(60, D) * 0.5, seeded_normal(61, D)])

    def f(z):
        q, p = z[:D][None], z[D:][None]
        st = orm.IntegratorState(q, p, p, torch.zeros(1, dtype=F64), vmap(grad(M.logp))(q))
        o = step(st, h)
        return torch.cat([o.position[0], o.momentum[0]])

    J = torch.autograd.functional.jacobian(f, z)
    assert abs(torch.linalg.det(J).item() - 1.0) < 1e-10


def test_energy_error_is_second_order():
    D, T = 3, 0.8
    M, _, ke = _monge_step(D, 1.0)
    q0, p0 = seeded_normal(70, 1, D) * 0.5, seeded_normal(71, 1, D)
    energy = orm.hmc_energy(ke)
    errs = []
    for n in (8, 16, 32):
        _, step, _ = _monge_step(D, 1.0)
        st = orm.IntegratorState(q0, p0, p0, vmap(M.logp)(q0), vmap(grad(M.logp))(q0))
        e0 = energy(st)
        for _ in range(n):
            st = step(st, T / n)
        errs.append(abs((energy(st) - e0).item()))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.0 < r1 < 5.5 and 3.0 < r2 < 5.5, errs


def test_divergent_and_nan_proposals_are_rejected():
    """Huge step: energy blows up -> is_divergent; NaN -> weight -inf, rejected (proposal.py:107-108)."""
    D = 5
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, 1.0, D)
    init_fn, step_fn = orm.rmhmc(M.logp, 3.0, metric, 5, max_iters=8)
    q0 = seeded_normal(80, 6, D)
    s0 = init_fn(q0)
    s1, info = step_fn((seeded_normal(81, 6, D), torch.full((6,), 0.5, dtype=F64)), s0)
    w = info.proposal.weight
    assert torch.all(~torch.isnan(w))
    bad = ~torch.isfinite(info.energy) | (-w > 1000)
    assert bad.any()
    assert torch.all(info.is_divergent[bad] | torch.isinf(w[bad]))
    assert torch.all(~info.is_accepted[bad])
    assert torch.equal(s1.position[bad], q0[bad])
