"""CPU pin of the derivative oracle for the LMC-Monge alpha2 tangent.

The GPU tangent is compared with torch.func.jacfwd through the oracle's
lmcmonge kernel (oracle/lmcmonge.py).  That derivative is only a pin if every
alpha2 dependency of the oracle is traced (the velocity draw, the metric's
normalizing constant, both half steps); a detached one (a float() of alpha2,
say) would make jacfwd silently drop a term.  Central differences of the
oracle itself catch that: they see every dependency.
"""
import torch
from torch.func import jacfwd

from oracle import lmcmonge as olm
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform


def _prop(D, C, eps, L, faithful, seed):
    M = om.NealFunnel(D)
    q0 = 0.5 * seeded_normal(seed, C, D)
    z, u = seeded_normal(seed + 1, C, D), seeded_uniform(seed + 2, C)
    im = 0.5 + seeded_uniform(seed + 3, D)
    kern = olm.build_kernel(reference_faithful=faithful)

    def prop(a2):
        _, info = kern((z, u), olm.init(q0, M.logp), M.logp, eps, im, L, a2)
        return info.proposal.state.position

    return prop


def test_jacfwd_through_the_oracle_equals_central_differences():
    for faithful in (True, False):
        prop = _prop(6, 3, 0.05, 6, faithful, 10)
        a2, h = 0.7, 1e-5
        J = jacfwd(prop)(torch.tensor(a2, dtype=torch.float64))
        fd = (prop(a2 + h) - prop(a2 - h)) / (2 * h)
        assert rel_err(J, fd) < 1e-7, faithful
        assert J.abs().max() > 1e-4  # alpha2 does move the proposal
