"""CPU pins of the NUTS oracle (oracle/nuts.py).

* one doubling (M = 1) is one implicit-midpoint step in the drawn direction
  followed by a Metropolis test with min(1, e^dH): it equals the RMHMC oracle
  with one integration step of +eps or -eps fed the same momentum and uniform
  (the biased merge against the start, weight 0, is that test);
* the numpyro checkpoint indices follow the reference's bit rules
  (termination.py:71-87: 6 -> (3, 2), 7 -> (0, 2), 13 -> (2, 2));
* trajectory bookkeeping: num_integration_steps <= 2^M - 1, the acceptance
  statistic is in [0, 1], and a trajectory stopped only by the doubling cap
  has exactly 2^M - 1 states.
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
