"""CPU pins of the NUTS oracle (oracle/nuts.py).

* one doubling (M = 1) is one implicit-midpoint step in the drawn direction
  followed by a Metropolis test with min(1, e^dH): it equals the RMHMC oracle
  with one integration step of +eps or -eps fed the same momentum and uniform
  (the biased merge against the start, weight 0, is that test);
* the numpyro checkpoint indices follow the reference's bit rules
  (termination.py:71-87: 6 -> (3, 2), 7 -> (0, 2), 13 -> (2, 2));
* trajectory bookkeeping: num_integration_steps <= 2^M - 1, the acceptance
  statistic is in [0, 1], and a trajectory stopped only by the doubling cap
  has exactly 2^M - 1 states;
* the U-turn criteria (geomjax/rmhmc/metrics.py:100-117): a truth table on
  vectors where each criterion has one answer; under the identity metric
  (momentum = velocity) the three give identical trajectories, under the
  Monge metric three different ones.
"""
import math

import torch

from oracle import models as om
from oracle import nuts as onuts
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal, seeded_uniform


def _funnel(D):
    M = om.NealFunnel(D)
    return M, om.monge_metric_fn(M.logp, 1.0, D)


def test_leaf_index_rules():
    assert onuts._leaf_idx_to_ckpt_idxs(6) == (3, 2)
    assert onuts._leaf_idx_to_ckpt_idxs(7) == (0, 2)
    assert onuts._leaf_idx_to_ckpt_idxs(13) == (2, 2)
    assert onuts._leaf_idx_to_ckpt_idxs(0) == (1, 0)


def test_one_doubling_is_one_rmhmc_step():
    D, C, eps = 5, 6, 0.1
    M, metric = _funnel(D)
    q0 = seeded_normal(1, C, D)
    z = seeded_normal(2, C, D)
    u = seeded_uniform(3, C)
    u_dir = torch.tensor([[0.2], [0.7], [0.4], [0.9], [0.1], [0.6]], dtype=torch.float64)
    s0 = orm.init(q0, M.logp)
    new, infos = onuts.build_kernel(max_iters=8)((z, u_dir, u[:, None], torch.zeros(C, 1, dtype=torch.float64)),
                                                 s0, M.logp, eps, metric, max_num_doublings=1)
    k = orm.build_kernel(max_iters=8)
    for c in range(C):
        sgn = 1.0 if float(u_dir[c, 0]) < 0.5 else -1.0
        sr, ir = k((z[c:c + 1], u[c:c + 1]), orm.RMHMCState(*(t[c:c + 1] for t in s0)), M.logp, sgn * eps, metric, 1)
        assert rel_err(new.position[c:c + 1], sr.position) < 1e-12, c
        assert abs(infos[c].acceptance_rate - float(ir.acceptance_rate[0])) < 1e-12, c
        assert infos[c].num_integration_steps == 1 and infos[c].num_trajectory_expansions == 1


def test_trajectory_bookkeeping():
    D, C, eps, Mx = 4, 3, 0.02, 3
    M, metric = _funnel(D)
    q0 = 0.5 * seeded_normal(4, C, D)
    noise = (seeded_normal(5, C, D), seeded_uniform(6, C, Mx), seeded_uniform(7, C, Mx),
             seeded_uniform(8, C, 2 ** Mx - 1))
    _, infos = onuts.build_kernel(max_iters=8)(noise, orm.init(q0, M.logp), M.logp, eps, metric, max_num_doublings=Mx)
    for info in infos:
        assert 1 <= info.num_integration_steps <= 2 ** Mx - 1
        assert 0.0 <= info.acceptance_rate <= 1.0 + 1e-12
        if not info.is_turning and not info.is_divergent:
            assert info.num_integration_steps == 2 ** Mx - 1 and info.num_trajectory_expansions == Mx


def test_criteria_definitions():
    """geomjax/rmhmc/metrics.py:100-117 on vectors where each criterion has one answer."""
    e1, m1 = torch.tensor([1.0, 0.0], dtype=torch.float64), torch.tensor([-1.0, 0.0], dtype=torch.float64)
    # momenta point back, velocities forward: only the criteria that read the momenta turn
    args = dict(m_left=m1, m_right=m1, m_sum=m1, v_left=e1, v_right=e1, v_sum=e1)
    assert onuts._is_turning("euc", **args) is False
    assert onuts._is_turning("bet", **args) is True    # v_left . m_sum < 0
    assert onuts._is_turning("riem", **args) is True   # m_left . v_sum < 0
    # only the right end's momentum points back: "riem" turns at the right end alone
    args = dict(m_left=e1, m_right=m1, m_sum=e1, v_left=e1, v_right=e1, v_sum=e1)
    assert [onuts._is_turning(c, **args) for c in ("euc", "bet", "riem")] == [False, False, True]
    # orthogonal counts as turning (<= 0)
    e2 = torch.tensor([0.0, 1.0], dtype=torch.float64)
    assert onuts._is_turning("euc", m_left=e1, m_right=e1, m_sum=e1, v_left=e2, v_right=e1, v_sum=e1) is True


def _run(D, alpha2, criterion, seed, C=4, Mx=5, eps=0.15):
    M = om.NealFunnel(D)
    metric = om.monge_metric_fn(M.logp, alpha2, D)
    q0 = 0.5 * seeded_normal(seed, C, D)
    noise = (seeded_normal(seed + 1, C, D), seeded_uniform(seed + 2, C, Mx), seeded_uniform(seed + 3, C, Mx),
             seeded_uniform(seed + 4, C, 2 ** Mx - 1))
    return onuts.build_kernel(max_iters=8)(noise, orm.init(q0, M.logp), M.logp, eps, metric, max_num_doublings=Mx,
                                           stopping_criterion=criterion)


def test_criteria_coincide_under_the_identity_metric():
    """With alpha2 = 0 the metric is I, momentum = velocity, and the three
    criteria are the same test: identical trajectories."""
    runs = [_run(4, 0.0, c, 20) for c in ("euc", "bet", "riem")]
    for s, infos in runs[1:]:
        assert torch.equal(s.position, runs[0][0].position)
        assert [i.num_integration_steps for i in infos] == [i.num_integration_steps for i in runs[0][1]]


def test_criteria_differ_under_the_monge_metric():
    """With alpha2 = 1 momentum and velocity differ, and so do the trajectory
    lengths the three criteria give on this seed."""
    steps = {c: tuple(i.num_integration_steps for i in _run(4, 1.0, c, 30)[1]) for c in ("euc", "bet", "riem")}
    assert len(set(steps.values())) == 3, steps


def test_lmcmonge_one_doubling_is_one_lmc_step():
    """NUTS-LMC-Monge (geomjax/lmcmonge/nuts.py) with M = 1 is one lan_integrator
    step of +-eps and the Metropolis test on -l + K - J: the lmcmonge oracle with
    one integration step of that signed size, fed the same velocity normals and
    the merge uniform as its accept uniform (non-unit mass, alpha2 = 0.6, a
    carried volume adjustment; both algebras)."""
    from oracle import lmcmonge as olm

    D, C, eps = 5, 6, 0.1
    M = om.NealFunnel(D)
    im = 0.5 + seeded_uniform(11, D)
    q0 = 0.5 * seeded_normal(12, C, D)
    z, u = seeded_normal(13, C, D), seeded_uniform(14, C)
    u_dir = torch.tensor([[0.2], [0.7], [0.4], [0.9], [0.1], [0.6]], dtype=torch.float64)
    for faithful in (True, False):
        s0 = olm.init(q0, M.logp)
        s0 = s0._replace(volume_adjustment=0.01 * seeded_normal(15, C))
        kern = onuts.build_lmcmonge_kernel(max_num_doublings=1, reference_faithful=faithful)
        new, infos = kern((z, u_dir, u[:, None], torch.zeros(C, 1, dtype=torch.float64)), s0, M.logp, eps, im, 0.6)
        k = olm.build_kernel(reference_faithful=faithful)
        for c in range(C):
            sgn = 1.0 if float(u_dir[c, 0]) < 0.5 else -1.0
            sc = olm.LMCState(*(t[c:c + 1] for t in s0))
            sr, ir = k((z[c:c + 1], u[c:c + 1]), sc, M.logp, sgn * eps, im, 1, 0.6)
            assert rel_err(new.position[c:c + 1], sr.position) < 1e-12, (faithful, c)
            assert abs(float(new.volume_adjustment[c]) - float(sr.volume_adjustment[0])) < 1e-12, (faithful, c)
            assert abs(infos[c].acceptance_rate - float(ir.acceptance_rate[0])) < 1e-12, (faithful, c)
            assert rel_err(infos[c].momentum, ir.velocity[0]) < 1e-14  # NUTSInfo carries the initial velocity


def test_lmcmonge_criteria_coincide_under_the_identity_metric():
    """alpha2 = 0 and unit mass: the LMC-Monge momentum M v + alpha2 L (v.g) g is
    the velocity, and the three criteria give identical trajectories."""
    from oracle import lmcmonge as olm

    D, C, Mx = 4, 4, 5
    M = om.NealFunnel(D)
    q0 = 0.5 * seeded_normal(40, C, D)
    noise = (seeded_normal(41, C, D), seeded_uniform(42, C, Mx), seeded_uniform(43, C, Mx),
             seeded_uniform(44, C, 2 ** Mx - 1))
    im = torch.ones(D, dtype=torch.float64)
    runs = []
    for crit in ("euc", "bet", "riem"):
        kern = onuts.build_lmcmonge_kernel(max_num_doublings=Mx)
        runs.append(kern(noise, olm.init(q0, M.logp), M.logp, 0.15, im, 0.0, stopping_criterion=crit))
    for s, infos in runs[1:]:
        assert torch.equal(s.position, runs[0][0].position)
        assert [i.num_integration_steps for i in infos] == [i.num_integration_steps for i in runs[0][1]]


def test_lmcmc_one_doubling_is_one_lmc_step():
    """NUTS-LMC (general metric, geomjax/lmcmc/nuts.py) with M = 1 is one
    lan_integrator step of +-eps with the Metropolis test on -l + K - J: the lmcmc
    oracle with one integration step of that signed size (logistic regression,
    Fisher metric, a carried volume adjustment)."""
    from oracle import lmcmc as olc
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(60, 4, 3)
    Mo = om.LogisticRegression(X, y, 100.0)
    C, eps = 5, 0.2
    q0 = 0.3 * seeded_normal(50, C, 4)
    z, u = seeded_normal(51, C, 4), seeded_uniform(52, C)
    u_dir = torch.tensor([[0.2], [0.7], [0.4], [0.9], [0.1]], dtype=torch.float64)
    s0 = olc.init(q0, Mo.logp)
    s0 = s0._replace(volume_adjustment=0.01 * seeded_normal(53, C))
    new, infos = onuts.build_lmcmc_kernel()((z, u_dir, u[:, None], torch.zeros(C, 1, dtype=torch.float64)), s0,
                                            Mo.logp, eps, Mo.fisher_metric_fn, max_num_doublings=1)
    k = olc.build_kernel()
    for c in range(C):
        sgn = 1.0 if float(u_dir[c, 0]) < 0.5 else -1.0
        sr, ir = k((z[c:c + 1], u[c:c + 1]), olc.LMCState(*(t[c:c + 1] for t in s0)), Mo.logp, sgn * eps,
                   Mo.fisher_metric_fn, 1)
        assert rel_err(new.position[c:c + 1], sr.position) < 1e-12, c
        assert abs(float(new.volume_adjustment[c]) - float(sr.volume_adjustment[0])) < 1e-12, c
        assert abs(infos[c].acceptance_rate - float(ir.acceptance_rate[0])) < 1e-12, c
        assert rel_err(infos[c].momentum, ir.velocity[0]) < 1e-14  # NUTSInfo carries the initial velocity


def _leaf_selection_worst_z():
    """Worst z-score of the empirical leaf frequencies against the closed form."""
    D, eps, M, trials = 3, 0.5, 3, 6000  # a step long enough that the eight weights differ
    Mdl, metric = _funnel(D)
    q0 = 0.3 * seeded_normal(66, 1, D)
    z = seeded_normal(67, 1, D)
    mg, ke, imvp = onuts.gaussian_riemannian(metric)
    raw_step = onuts.implicit_midpoint(Mdl.logp, ke, imvp, max_iters=20)
    energy_fn = onuts.hmc_energy(ke)
    momentum = torch.func.vmap(mg)(z, q0)
    velocity = torch.func.vmap(imvp)(q0, momentum)
    s0 = orm.init(q0, Mdl.logp)
    s0 = onuts.IntegratorState(q0, momentum, velocity, s0.logdensity, s0.logdensity_grad)
    cache, ecache = {}, {}

    def one_step(state, step):  # the trajectory is fixed: integrate each (state, direction) once
        key = (id(state), step)
        if key not in cache:
            cache[key] = raw_step(state, step)
        return cache[key]

    def energy(state):
        if id(state) not in ecache:
            ecache[id(state)] = energy_fn(state)
        return ecache[id(state)]

    H0 = float(energy(s0)[0])
    # the leaves in time order: subtree 1 = one step forward, 2 = two back, 3 = four forward
    fwd, bwd = [s0], [s0]
    for _ in range(5):
        fwd.append(one_step(fwd[-1], eps))
    for _ in range(2):
        bwd.append(one_step(bwd[-1], -eps))
    subtrees = [fwd[1:2], bwd[1:3], fwd[2:6]]
    w = lambda s: math.exp(H0 - float(energy(s)[0]))
    W_prev = w(s0)
    a = []
    for S in subtrees:
        WS = sum(w(s) for s in S)
        a.append(min(1.0, WS / W_prev))
        W_prev += WS
    probs = {id(s0): math.prod(1 - x for x in a)}
    for j, S in enumerate(subtrees):
        WS = sum(w(s) for s in S)
        keep = a[j] * math.prod(1 - x for x in a[j + 1:])
        for s in S:
            probs[id(s)] = w(s) / WS * keep
    assert abs(sum(probs.values()) - 1.0) < 1e-12
    assert max(a) < 1.0 and min(a) < 0.9  # every merge is a real coin, not a sure keep
    leaves = {id(s): s for S in subtrees for s in S} | {id(s0): s0}

    u_dir = torch.tensor([0.2, 0.8, 0.3], dtype=torch.float64)
    g = torch.Generator().manual_seed(62)
    counts = {k: 0 for k in probs}
    for _ in range(trials):
        u_merge = torch.rand(M, generator=g, dtype=torch.float64)
        u_step = torch.rand(2 ** M - 1, generator=g, dtype=torch.float64)
        prop, traj, e, div, turn = onuts._expand(s0, H0, one_step, energy, eps, u_dir, u_merge, u_step, M, "euc",
                                                  1000.0)
        assert e == M and not div and not turn and traj.num_states == 2 ** M - 1  # the fixed trajectory
        hit = [k for k, s in leaves.items() if torch.equal(prop.state.position, s.position)]
        assert len(hit) == 1
        counts[hit[0]] += 1
    worst = 0.0
    for k, p in probs.items():
        se = math.sqrt(p * (1 - p) / trials)
        zval = abs(counts[k] / trials - p) / se
        worst = max(worst, zval)
    print("merge probabilities", [round(x, 3) for x in a], "leaf probabilities",
          sorted(round(p, 4) for p in probs.values()), "worst z", round(worst, 2))
    return worst


def test_multi_doubling_leaf_selection_matches_the_closed_form():
    """With the directions fixed (+, -, +) and no U-turn or divergence within
    three doublings, the trajectory's eight states are fixed; only the merge
    and in-subtree uniforms vary.  Progressive uniform sampling inside subtree
    S_j picks leaf i with probability w_i / W(S_j), and the biased merge keeps
    S_j's pick with probability a_j = min(1, W(S_j) / W_{j-1}) (W_{j-1}: the
    total weight of the trajectory before the merge), so the transition ends
    at leaf i of S_j with probability (w_i / W(S_j)) a_j prod_{k>j} (1 - a_k)
    and at the start with prod_k (1 - a_k).  The empirical frequencies of the
    oracle's expansion over many uniforms must match (5 standard errors)."""
    assert _leaf_selection_worst_z() < 5.0


def test_an_unbiased_merge_is_caught(monkeypatch):
    """Planted mistake: the trajectory-level merge uses the progressive uniform
    rule instead of the biased one."""
    orig = onuts._sample
    monkeypatch.setattr(onuts, "_sample", lambda u, proposal, new, biased: orig(u, proposal, new, False))
    assert _leaf_selection_worst_z() > 5.0
