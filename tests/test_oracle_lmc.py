"""Pins of the oracle's LMC restatements (lmcmonge and lmcmc), CPU float64.

The two restatements compute the same Lagrangian step in independent ways:
lmcmonge through the reference's closed-form rank-one algebra and HVPs,
lmcmc through the full D x D x D jacfwd of the metric, three einsums and LU
factorisations.  With the Monge metric_fn and alpha2 = 1 (where the
reference's missing alpha2 is a factor 1) they must agree on positions,
velocities, volume adjustment and energy.  The volume adjustment is further
pinned by the Jacobian determinant of the step map, computed by autodiff.
"""
import pytest
import torch
from torch.func import grad, vmap

from oracle import lmcmc as olc
from oracle import lmcmonge as olm
from oracle import mcmc as omc
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

F64 = torch.float64


def test_identity_metric_lmc_matches_leapfrog_hmc():
    """tests/test_samplers.py:34-39,57: lmc(G=I) == hmc (rtol 1e-4)."""
    D = 2
    M = om.NealFunnel(D)
    z, u = seeded_normal(0, 3, D), seeded_uniform(1, 3)
    q0 = torch.zeros(3, D, dtype=F64)
    l1, _ = olc.build_kernel()((z, u), olc.init(q0, M.logp), M.logp, 1e-2, om.identity_metric_fn(D), 10)
    h1, _ = omc.hmc_kernel((z, u), omc.hmc_init(q0, M.logp), M.logp, 1e-2, torch.ones(D, dtype=F64), 10)
    assert torch.allclose(l1.position, h1.position, rtol=1e-4)


def test_monge_alpha0_quirk():
    """tests/test_samplers.py:58-59 leave the Monge comparison commented out: with
    alpha2 = 0 the reference's half step still applies its Sherman-Morrison
    correction (the alpha2 factor is missing, integrators.py:186-189), so it
    does NOT reduce to leapfrog HMC; with the factor restored it does."""
    D = 2
    M = om.NealFunnel(D)
    z, u = seeded_normal(2, 3, D), torch.zeros(3, dtype=F64)
    q0 = torch.zeros(3, D, dtype=F64)
    h1, _ = omc.hmc_kernel((z, u), omc.hmc_init(q0, M.logp), M.logp, 1e-2, torch.ones(D, dtype=F64), 10)
    m_ref, _ = olm.build_kernel()((z, u), olm.init(q0, M.logp), M.logp, 1e-2, torch.ones(D, dtype=F64), 10, 0.0)
    m_fix, _ = olm.build_kernel(reference_faithful=False)((z, u), olm.init(q0, M.logp), M.logp, 1e-2,
                                                          torch.ones(D, dtype=F64), 10, 0.0)
    assert not torch.allclose(m_ref.position, h1.position, rtol=1e-4)
    assert torch.allclose(m_fix.position, h1.position, rtol=1e-12, atol=1e-14)


def _monge_states(D, q0, v0, alpha2, faithful):
    M = om.NealFunnel(D)
    im = torch.ones(D, dtype=F64)
    (vg, ke, swg, nc, mvp) = olm.gaussian_riemannian(alpha2, im, faithful)
    g, l = vmap(torch.func.grad_and_value(M.logp))(q0)
    Lm = nc(alpha2, g)
    sdm = torch.sqrt(Lm)[:, None]
    gn = g / sdm
    hvp = vmap(lambda x, v: olm.hvp(M.logp, x, v))
    dl = swg(gn)
    Hdl = hvp(q0, dl) / sdm
    st = olm.RiemannianIntegratorState(torch.tensor(alpha2, dtype=F64), q0, mvp(v0, alpha2, g, Lm), v0, l, gn, dl,
                                       Hdl, swg(Hdl), hvp(q0, v0) / sdm, Lm, torch.zeros(q0.shape[0], dtype=F64))
    step = olm.lan_integrator(M.logp, swg, nc, mvp, faithful)
    return M, st, step, olm.lmcmonge_energy(ke)


@pytest.mark.parametrize("alpha2,faithful", [(1.0, True), (0.4, False)])
def test_lmcmonge_equals_lmcmc_with_monge_metric(alpha2, faithful):
    D, C, eps, n = 4, 3, 0.05, 6
    q0, v0 = seeded_normal(10, C, D) * 0.6, seeded_normal(11, C, D) * 0.5
    M, sm, step_m, energy_m = _monge_states(D, q0, v0, alpha2, faithful)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    vg, ke, omt, gld, mvp = olc.gaussian_riemannian(metric)
    g, l = vmap(torch.func.grad_and_value(M.logp))(q0)
    sc = olc.IntegratorState(q0, vmap(mvp)(q0, v0), v0, l, g, torch.zeros(C, dtype=F64))
    step_c = olc.lan_integrator(M.logp, omt, gld, mvp)
    energy_c = olc.lmc_energy(ke)
    assert rel_err(energy_m(sm), energy_c(sc)) < 1e-12
    for _ in range(n):
        sm, sc = step_m(sm, eps), step_c(sc, eps)
    assert rel_err(sm.position, sc.position) < 1e-11
    assert rel_err(sm.velocity, sc.velocity) < 1e-11
    assert rel_err(sm.volume_adjustment, sc.volume_adjustment) < 1e-9
    assert rel_err(energy_m(sm), energy_c(sc)) < 1e-11
    if not faithful:  # exact momentum G v
        assert rel_err(sm.momentum, sc.momentum) < 1e-11


def test_reference_lmcmonge_differs_from_lmcmc_when_alpha2_not_one():
    """The missing alpha2 matters off alpha2 = 1 (documents the quirk)."""
    D, C, eps = 4, 3, 0.05
    q0, v0 = seeded_normal(12, C, D) * 0.6, seeded_normal(13, C, D) * 0.5
    _, sm, step_m, _ = _monge_states(D, q0, v0, 0.4, True)
    _, sx, step_x, _ = _monge_states(D, q0, v0, 0.4, False)
    for _ in range(4):
        sm, sx = step_m(sm, eps), step_x(sx, eps)
    assert rel_err(sm.position, sx.position) > 1e-8


def _jac_logdet(step_fn, make_state, q, v):
    D = q.shape[0]

    def f(z):
        st = make_state(z[:D][None], z[D:][None])
        o = step_fn(st)
        return torch.cat([o.position[0], o.velocity[0]])

    J = torch.autograd.functional.jacobian(f, torch.cat([q, v]))
    return torch.linalg.slogdet(J)[1]


def test_lmcmc_volume_adjustment_is_log_jacobian():
    """J1 - J0 equals log|det d(q1, v1)/d(q0, v0)| (Lagrangian steps are not volume preserving)."""
    D, eps = 3, 0.2
    M = om.NealFunnel(D)
    metric = M.fisher_metric_fn
    vg, ke, omt, gld, mvp = olc.gaussian_riemannian(metric)
    step = olc.lan_integrator(M.logp, omt, gld, mvp)

    def make(q, v):
        g, l = vmap(torch.func.grad_and_value(M.logp))(q)
        return olc.IntegratorState(q, vmap(mvp)(q, v), v, l, g, torch.zeros(1, dtype=F64))

    q, v = seeded_normal(20, D) * 0.5, seeded_normal(21, D) * 0.5
    ld = _jac_logdet(lambda s: step(s, eps), make, q, v)
    dJ = step(make(q[None], v[None]), eps).volume_adjustment[0]
    assert abs(ld - dJ) < 1e-9


@pytest.mark.parametrize("alpha2,faithful", [(1.0, True), (0.6, False)])
def test_lmcmonge_volume_adjustment_is_log_jacobian(alpha2, faithful):
    D, eps = 3, 0.15
    q, v = seeded_normal(22, D) * 0.5, seeded_normal(23, D) * 0.5

    def make(qb, vb):
        return _monge_states(D, qb, vb, alpha2, faithful)[1]

    step = _monge_states(D, q[None], v[None], alpha2, faithful)[2]
    ld = _jac_logdet(lambda s: step(s, eps), make, q, v)
    dJ = step(make(q[None], v[None]), eps).volume_adjustment[0]
    assert abs(ld - dJ) < 1e-9


def test_lmcmonge_velocity_draw_covariance():
    """v0 = chol(M^-1 - alpha2 dl dl^T) z has covariance G^-1 with G = M + alpha2 g g^T."""
    D, alpha2 = 5, 0.7
    M = om.NealFunnel(D)
    im = 0.5 + seeded_uniform(30, D)
    vg, _, swg, nc, _ = olm.gaussian_riemannian(alpha2, im)
    q = seeded_normal(31, 1, D)
    g = vmap(grad(M.logp))(q)
    dl = swg(g / torch.sqrt(nc(alpha2, g))[:, None])
    cols = torch.stack([vg(e[None], q, alpha2, dl)[0] for e in torch.eye(D, dtype=F64)], dim=1)
    G = torch.diag(1 / im) + alpha2 * torch.outer(g[0], g[0])
    assert rel_err(cols @ cols.T, torch.linalg.inv(G)) < 1e-12
