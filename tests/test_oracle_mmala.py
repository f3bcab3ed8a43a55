"""Pins of the oracle's manifold MALA (geomjax/lmcmc/mmala.py), CPU float64.

* Gamma (jacfwd of inv(metric_fn), einsum "ijj") against closed forms: a
  diagonal metric exp(x) (Gamma = -exp(-x)), a 1-D metric 1 + x^2, and the
  logistic Fisher metric (Gamma = -G^-1 X^T (w * a), w = s(1-s)(1-2s),
  a_n = x_n^T G^-1 x_n).
* With the identity metric the step is Euclidean MALA: the position is
  x + eps g + sqrt(2 eps) z and the weight is the textbook log Metropolis-Hastings
  ratio of MALA, negated in the reference's pairing.
* With a constant metric M the draw is mu + chol(2 eps M^-1) z with the lower
  Cholesky factor of the covariance (scipy's routine).
"""
import math

import numpy as np
import scipy.linalg
import torch

from oracle import mmala as omm
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

F64 = torch.float64


def test_gamma_closed_forms():
    x = seeded_normal(1, 5)
    g = omm.get_Gamma(x, lambda th: torch.diag(torch.exp(th)))
    assert rel_err(g, -torch.exp(-x)) < 1e-14
    x1 = torch.tensor([0.7], dtype=F64)
    g1 = omm.get_Gamma(x1, lambda th: (1.0 + th * th).reshape(1, 1))
    assert abs(float(g1[0]) - (-2 * 0.7 / (1 + 0.49) ** 2)) < 1e-14


def test_gamma_logistic_fisher_closed_form():
    from geomjax_amd.synthetic import make_logistic

    X, y, _ = make_logistic(80, 6, 2)
    M = om.LogisticRegression(X, y, 100.0)
    beta = 0.3 * seeded_normal(3, 6)
    s = torch.sigmoid(M.X @ beta)
    w = s * (1 - s) * (1 - 2 * s)
    G = M.fisher_metric_fn(beta)
    Gi = torch.linalg.inv(G)
    a = torch.einsum("nd,de,ne->n", M.X, Gi, M.X)
    closed = -Gi @ (M.X.T @ (w * a))
    assert rel_err(omm.get_Gamma(beta, M.fisher_metric_fn), closed) < 1e-12


def _gauss_logp(x):
    return -0.5 * torch.dot(x, x)


def test_identity_metric_is_mala_with_the_reference_sign():
    D, C, eps = 3, 4, 0.2
    eye = lambda th: torch.eye(D, dtype=F64)
    q0 = seeded_normal(4, C, D)
    z, u = seeded_normal(5, C, D), seeded_uniform(6, C)
    s0 = omm.init(q0, _gauss_logp, eye)
    assert torch.count_nonzero(s0.Gamma) == 0
    for faithful in (True, False):
        new, info = omm.build_kernel(reference_faithful=faithful)((z, u), s0, _gauss_logp, eps, eye)
        for c in range(C):
            x, n = q0[c], z[c]
            x1 = x + eps * (-x) + math.sqrt(2 * eps) * n
            assert rel_err(info.proposal.position[c], x1) < 1e-14
            # textbook MALA: log pi(x') + log q(x | x') - log pi(x) - log q(x' | x)
            fwd = -torch.sum((x1 - x - eps * (-x)) ** 2) / (4 * eps)
            bwd = -torch.sum((x - x1 - eps * (-x1)) ** 2) / (4 * eps)
            log_alpha = float(_gauss_logp(x1) + bwd - _gauss_logp(x) - fwd)
            want = -log_alpha if faithful else log_alpha
            assert abs(float(info.weight[c]) - want) < 1e-12, (faithful, c)
            p = min(1.0, math.exp(want))
            assert abs(float(info.acceptance_rate[c]) - p) < 1e-14
            assert bool(info.is_accepted[c]) == (float(u[c]) < p)
            assert rel_err(new.position[c], x1 if float(u[c]) < p else x) == 0.0


def test_constant_metric_draw_uses_the_lower_cholesky_of_the_covariance():
    D, eps = 4, 0.1
    A = seeded_normal(7, D, D)
    Mm = A @ A.T + D * torch.eye(D, dtype=F64)
    metric = lambda th: Mm
    q0 = seeded_normal(8, 1, D)
    z = seeded_normal(9, 1, D)
    s0 = omm.init(q0, _gauss_logp, metric)
    assert torch.count_nonzero(s0.Gamma) == 0
    _, info = omm.build_kernel()((z, torch.zeros(1, dtype=F64)), s0, _gauss_logp, eps, metric)
    Mn = Mm.numpy()
    mu = q0[0].numpy() + eps * np.linalg.solve(Mn, -q0[0].numpy())
    Lc = scipy.linalg.cholesky(2 * eps * np.linalg.inv(Mn), lower=True)
    want = mu + Lc @ z[0].numpy()
    assert rel_err(info.proposal.position[0], torch.from_numpy(want)) < 1e-13
