"""CPU: the ChEES driver's building blocks (geomjax_amd.adaptation).

* Halton jitter: the base-2 van der Corput values of i + 1;
* Adam: iterate-for-iterate equal to torch.optim.Adam (same conventions as
  optax.adam: bias-corrected moments, eps outside the square root);
* dual averaging: converges to the root of a monotone acceptance curve
  (a sign or scale slip does not);
* ChEES trajectory gradient: equals autograd of
  (|q_L - m_L|^2 - |q_0 - m_0|^2)^2 / 4 in log T along straight trajectories
  q_L = q_0 + jitter T v (means held fixed), acceptance-weighted over the
  non-divergent chains.
"""
import math

import torch

from geomjax_amd.adaptation import adam, chees_trajectory_gradient, dual_averaging, halton, harmonic_acceptance


def test_halton_is_van_der_corput_base_2():
    assert [halton(i) for i in range(7)] == [0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]
    t = halton(torch.arange(40), 10)
    assert torch.allclose(t, torch.tensor([halton(i, 10) for i in range(40)], dtype=torch.float64), rtol=0, atol=0)


def test_adam_matches_torch_optim_adam():
    opt = adam(0.1)
    x, st = 0.0, opt.init(0.0)
    p = torch.zeros((), dtype=torch.float64, requires_grad=True)
    ref = torch.optim.Adam([p], lr=0.1, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(60):
        upd, st = opt.update(2.0 * (x - 3.0), st, x)
        x += upd
        ref.zero_grad()
        ((p - 3.0) ** 2).backward()
        ref.step()
        assert abs(x - float(p)) <= 1e-12 * max(1.0, abs(x))


def test_harmonic_acceptance_skips_divergent_chains():
    acc = torch.tensor([0.5, 1.0, 0.25, 0.9])
    div = torch.tensor([False, False, False, True])
    # 3 / (1/0.5 + 1/1 + 1/0.25); the divergent chain's 0.9 is left out
    assert math.isclose(harmonic_acceptance(acc, div), 3.0 / 7.0, rel_tol=1e-15)
    assert math.isnan(harmonic_acceptance(acc, torch.ones(4, dtype=torch.bool)))
    # one chain that never accepts pulls the harmonic mean to 0
    assert harmonic_acceptance(torch.tensor([0.0, 1.0]), torch.zeros(2, dtype=torch.bool)) == 0.0


def test_dual_averaging_finds_the_target_acceptance():
    init, update, final = dual_averaging()
    target = 0.651
    st = init(1.0)
    for _ in range(3000):
        acc = math.exp(-math.exp(st.log_x))  # acceptance falls with the step size
        st = update(st, target - acc)
    assert abs(final(st) - (-math.log(target))) < 5e-3


def test_chees_trajectory_gradient_is_the_log_T_derivative():
    g = torch.Generator().manual_seed(3)
    C, D = 7, 5
    q0 = torch.randn(C, D, generator=g, dtype=torch.float64)
    v = torch.randn(C, D, generator=g, dtype=torch.float64)
    acc = torch.rand(C, generator=g, dtype=torch.float64)
    div = torch.zeros(C, dtype=torch.bool)
    div[2] = True
    jit, T = 0.7, 1.3
    qL = q0 + jit * T * v
    mL, m0 = qL.mean(0), q0.mean(0)
    logT = torch.tensor(math.log(T), dtype=torch.float64, requires_grad=True)
    crit = 0.25 * (((q0 + jit * torch.exp(logT) * v - mL) ** 2).sum(-1) - ((q0 - m0) ** 2).sum(-1)) ** 2
    grads = torch.stack([torch.autograd.grad(c, logT, retain_graph=True)[0] for c in crit])
    ok = ~div
    expect = float((acc * grads * ok).sum() / (acc * ok).sum())
    got = chees_trajectory_gradient(jit, T, qL, v, q0, acc, div)
    assert abs(got - expect) <= 1e-12 * max(1.0, abs(expect))


def test_chees_update_sign_of_the_trajectory_step():
    """The faithful update feeds the trajectory gradient to Adam as is (a positive
    gradient shortens the trajectory); reference_faithful=False lengthens it."""
    from geomjax_amd.adaptation import chees_base

    g = torch.Generator().manual_seed(5)
    C, D = 16, 4
    q0 = torch.randn(C, D, generator=g, dtype=torch.float64)
    v = torch.randn(C, D, generator=g, dtype=torch.float64)
    qL = q0 + 0.5 * v
    acc = torch.full((C,), 0.7, dtype=torch.float64)
    div = torch.zeros(C, dtype=torch.bool)
    grad = chees_trajectory_gradient(1.0, 0.5, qL, v, q0, acc, div)
    assert grad != 0.0
    moved = {}
    for faithful in (True, False):
        init_fn, update_fn = chees_base(lambda i: 1.0, lambda i: i + 1, adam(0.1), 0.651, 0.5, faithful)
        st = init_fn(0, 0.5)
        new = update_fn(st, qL, v, q0, acc, div)
        moved[faithful] = new.log_trajectory_length_moving_average - math.log(0.5)
    assert math.copysign(1.0, moved[True]) == -math.copysign(1.0, grad)
    assert math.copysign(1.0, moved[False]) == math.copysign(1.0, grad)


def test_chees_lmcmonge_alpha2_update():
    """chees_adaptation_lmcmonge.py:59-341: the faithful state starts at alpha2 =
    1.0 whatever the caller gives (:282-284) and, with update_alpha, moves log
    alpha2 against the alpha2 gradient (Adam on the gradient as computed); the
    gradient is the trajectory gradient's formula with dq/dalpha2 in place of the
    velocity; the step size and trajectory length updates equal the RMHMC
    driver's; without update_alpha alpha2 does not move."""
    from geomjax_amd.adaptation import chees_base, chees_lmcmonge_base

    g = torch.Generator().manual_seed(6)
    C, D = 16, 4
    q0 = torch.randn(C, D, generator=g, dtype=torch.float64)
    v = torch.randn(C, D, generator=g, dtype=torch.float64)
    dq = torch.randn(C, D, generator=g, dtype=torch.float64)
    qL = q0 + 0.5 * v
    acc = torch.full((C,), 0.7, dtype=torch.float64)
    div = torch.zeros(C, dtype=torch.bool)
    ga = chees_trajectory_gradient(1.0, 0.5, qL, dq, q0, acc, div)
    assert ga != 0.0
    for faithful in (True, False):
        init_fn, update_fn = chees_lmcmonge_base(lambda i: 1.0, lambda i: i + 1, adam(0.1), 0.651, 0.5,
                                                 update_alpha=True, reference_faithful=faithful)
        st = init_fn(0, 0.5, 0.01)
        assert st.alpha2 == (1.0 if faithful else 0.01)
        new = update_fn(st, qL, v, dq, q0, acc, div)
        moved = new.log_alpha2_moving_average - math.log(st.alpha2)
        # step weight 1 ** -0.5 = 1: the moving average is the new log alpha2; Adam's
        # first step has size lr = 0.1 whatever the gradient's scale
        assert abs(abs(moved) - 0.1) < 1e-6
        assert math.copysign(1.0, moved) == (-1.0 if faithful else 1.0) * math.copysign(1.0, ga)
        assert math.isclose(new.alpha2, math.exp(new.log_alpha2_moving_average))
        rm_init, rm_update = chees_base(lambda i: 1.0, lambda i: i + 1, adam(0.1), 0.651, 0.5, faithful)
        rm = rm_update(rm_init(0, 0.5), qL, v, q0, acc, div)
        assert tuple(new[:8]) == tuple(rm)
    init_fn, update_fn = chees_lmcmonge_base(lambda i: 1.0, lambda i: i + 1, adam(0.1), 0.651, 0.5)
    st = init_fn(0, 0.5, 0.01)
    assert update_fn(st, qL, v, dq, q0, acc, div).alpha2 == 1.0
