"""The glmc oracle (generalized LMC: persistent velocity, non-reversible slice)
pinned without retyping its formulas:

* with alpha = 1 the velocity is a fresh draw, and a slice that lands on u after
  its translation makes the slice test the Metropolis test with uniform u:
  the transition must equal one lmcmc step (the lmcmc oracle, itself pinned
  in test_oracle_lmc.py) -- position, log density, gradient, volume
  adjustment and accept decision; the returned velocity is the end velocity
  (accepted) or the negated initial one (rejected), and the returned slice
  u e^{-dE} (accepted) or u;
* with alpha = 0 the velocity is never refreshed: an accepted step continues
  from the integrator's end velocity, a rejected one reverses the initial one;
* on a standard Gaussian target with G = I, chains started from exact draws
  of (q, v, slice) stay exact (Neal's non-reversible slice update leaves the
  joint law invariant), and a slice that moves with e^{+dE} instead of
  e^{-dE} is caught by the same check.
"""
import torch

from oracle import glmc as og
from oracle import lmcmc as olc
from oracle import models as om
from oracle import mcmc as omc
from tests.helpers import rel_err, seeded_normal, seeded_uniform


def _funnel_monge(D, alpha2=1.0):
    M = om.NealFunnel(D)
    return M.logp, om.monge_metric_fn(M.logp, alpha2, D)


def _logistic(N, D, seed):
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(N, D, seed)
    M = om.LogisticRegression(X, y, 100.0)
    return M.logp, M.fisher_metric_fn


def test_alpha_one_equals_one_lmcmc_step():
    for logp, metric, D, eps, seed in [(*_funnel_monge(5), 5, 0.2, 1), (*_logistic(80, 4, 3), 4, 0.4, 2)]:
        C = 6
        q0 = 0.5 * seeded_normal(10 + seed, C, D)
        z = seeded_normal(20 + seed, C, D)
        u = seeded_uniform(30 + seed, C)
        delta = 0.3
        s_prev = torch.remainder(u - delta + 1.0, 2.0) - 1.0  # translates to exactly u
        so = olc.init(q0, logp)
        st = og.GLMCState(q0, seeded_normal(40 + seed, C, D), so.logdensity, so.logdensity_grad, s_prev,
                          so.volume_adjustment)
        g1, gi = og.build_kernel()((z, torch.zeros(C, dtype=torch.float64)), st, logp, eps, metric, 1.0, delta)
        l1, li = olc.build_kernel()((z, u), so, logp, eps, metric, 1)
        assert torch.equal(gi.is_accepted, li.is_accepted)
        for f in ("position", "logdensity", "logdensity_grad", "volume_adjustment"):
            assert rel_err(getattr(g1, f), getattr(l1, f)) < 1e-13, f
        v_end = -li.proposal.state.velocity  # lmcmc's proposal carries the flipped end velocity
        v0 = li.velocity
        expect_v = torch.where(li.is_accepted[:, None], v_end, -v0)
        assert rel_err(g1.velocity, expect_v) < 1e-13
        expect_s = torch.where(li.is_accepted, u * torch.exp(-li.proposal.weight), u)
        assert rel_err(g1.slice, expect_s) < 1e-13
        assert rel_err(gi.energy, li.energy) < 1e-13


def test_alpha_zero_keeps_the_velocity():
    logp, metric = _funnel_monge(4)
    C, D, eps = 8, 4, 0.15
    q0 = 0.4 * seeded_normal(50, C, D)
    so = olc.init(q0, logp)
    v_prev = 0.3 * seeded_normal(51, C, D)
    st = og.GLMCState(q0, v_prev, so.logdensity, so.logdensity_grad, 2 * seeded_uniform(52, C) - 1,
                      so.volume_adjustment)
    g1, gi = og.build_kernel()((seeded_normal(53, C, D), torch.zeros(C, dtype=torch.float64)), st, logp, eps,
                               metric, 0.0, 0.5)
    # one lan_integrator step from (q0, v_prev), as the lmcmc oracle takes it
    vg, ke, om_fn, gld, mvp = olc.gaussian_riemannian(metric)
    step = olc.lan_integrator(logp, om_fn, gld, mvp)
    mom = torch.func.vmap(mvp)(q0, v_prev)
    end = step(olc.IntegratorState(q0, mom, v_prev, so.logdensity, so.logdensity_grad, so.volume_adjustment), eps)
    acc = gi.is_accepted[:, None]
    assert rel_err(g1.velocity, torch.where(acc, end.velocity, -v_prev)) < 1e-13
    assert rel_err(g1.position, torch.where(acc, end.position, q0)) < 1e-13
    assert rel_err(gi.velocity, v_prev) == 0.0  # the info's velocity is the (unrefreshed) initial one


def _gaussian(q):
    return -0.5 * torch.dot(q, q)


def _identity(q):
    return torch.eye(q.shape[0], dtype=q.dtype)


def _stationary_run(kernel, C=4000, D=2, T=30, alpha=0.3, delta=0.37, eps=1.4, seed=7):
    g = torch.Generator().manual_seed(seed)
    q = torch.randn(C, D, generator=g, dtype=torch.float64)
    v = torch.randn(C, D, generator=g, dtype=torch.float64)
    s = 2 * torch.rand(C, generator=g, dtype=torch.float64) - 1
    so = olc.init(q, _gaussian)
    st = og.GLMCState(q, v, so.logdensity, so.logdensity_grad, s, so.volume_adjustment)
    acc = 0.0
    for _ in range(T):
        z = torch.randn(C, D, generator=g, dtype=torch.float64)
        st, info = kernel((z, torch.zeros(C, dtype=torch.float64)), st, _gaussian, eps, _identity, alpha, delta)
        acc += float(info.is_accepted.double().mean()) / T
    zs = []
    for x in (st.position, st.velocity):
        m, var = x.mean(0), x.var(0)
        zs.append(float((m.abs() / (1.0 / C) ** 0.5).max()))
        zs.append(float(((var - 1.0).abs() / (2.0 / C) ** 0.5).max()))
    us = torch.abs(st.slice)  # |slice| stays uniform on [0, 1]
    zs.append(abs(float(us.mean()) - 0.5) / (1.0 / 12.0 / C) ** 0.5)
    return max(zs), acc


def test_gaussian_target_stays_exact():
    zmax, acc = _stationary_run(og.build_kernel())
    print(f"max z {zmax:.2f}, acceptance {acc:.3f}")
    assert 0.3 < acc < 0.95  # a step long enough that energy errors matter (leapfrog is stable below 2)
    assert zmax < 5.0


def test_a_wrong_slice_update_is_caught(monkeypatch):
    """Planted mistake: on acceptance the slice moves with e^{+dE}."""
    def wrong(slice, proposal, new_proposal):
        d = new_proposal.weight
        do_accept = torch.log(torch.abs(slice)) <= d
        return omc._select(do_accept, new_proposal, proposal), do_accept, torch.where(
            do_accept, slice * torch.exp(d), slice)

    monkeypatch.setattr(og, "nonreversible_slice_sampling", wrong)
    zmax, _ = _stationary_run(og.build_kernel())
    print(f"planted mistake: max z {zmax:.2f}")
    assert zmax > 5.0
