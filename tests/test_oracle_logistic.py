"""Pins of the oracle's logistic-regression pieces against independent computations.

* log density: scipy's Bernoulli-logit and normal log densities;
* Fisher metric: the negative Hessian of the log density by autodiff (exact
  for the logit link: the Hessian does not depend on y);
* RMHMC kinetic energy -log N(p; 0, G(q)) (oracle: Cholesky + autodiff): its
  q-gradient equals the closed form 1/2 X^T [w (a - b^2)] (a_n = x_n^T G^-1 x_n,
  b = X G^-1 p, w = s(1-s)(1-2s)) and central finite differences; its
  p-gradient is G^-1 p;
* LMC: the oracle's Christoffel contraction (jacfwd of the metric) equals
  1/2 X^T diag(w (X v)) X, and its grad log det G equals X^T (w a).
A dropped term, a wrong sign or a transposed operand in the oracle fails one
of these.
"""
import numpy as np
import scipy.special
import scipy.stats
import torch
from torch.func import grad, hessian

from geomjax_amd.synthetic import make_logistic
from oracle import lmcmc as olm
from oracle import models as om
from oracle import rmhmc as orm
from tests.helpers import rel_err, seeded_normal


def _problem(N=120, D=7, seed=3):
    X, y, _ = make_logistic(N, D, seed)
    return om.LogisticRegression(X, y, 100.0), torch.as_tensor(X), torch.as_tensor(y)


def _closed_forms(X, q, p_or_v, prior=100.0):
    s = torch.sigmoid(X @ q)
    lam = s * (1 - s)
    w = lam * (1 - 2 * s)
    G = (X.T * lam) @ X + torch.eye(X.shape[1], dtype=torch.float64) / prior
    Gi = torch.linalg.inv(G)
    a = torch.einsum("ni,ij,nj->n", X, Gi, X)
    return s, lam, w, G, Gi, a


def test_logp_matches_scipy_bernoulli_and_normal():
    M, X, y = _problem()
    for seed in range(3):
        q = seeded_normal(10 + seed, X.shape[1])
        eta = (X @ q).numpy()
        ll = np.sum(y.numpy() * scipy.special.log_expit(eta) + (1 - y.numpy()) * scipy.special.log_expit(-eta))
        lp = scipy.stats.norm.logpdf(q.numpy(), 0.0, np.sqrt(100.0)).sum()
        assert abs(M.logp(q).item() - (ll + lp)) < 1e-10 * max(1.0, abs(ll + lp))


def test_fisher_metric_is_negative_hessian():
    M, X, _ = _problem()
    for seed in range(3):
        q = seeded_normal(20 + seed, X.shape[1])
        assert rel_err(M.fisher_metric_fn(q), -hessian(M.logp)(q)) < 1e-12


def test_rmhmc_kinetic_gradients_closed_form_and_fd():
    M, X, _ = _problem()
    _, kinetic_energy, inv_metric_vp = orm.gaussian_riemannian(M.fisher_metric_fn)
    for seed in range(3):
        q, p = seeded_normal(30 + seed, X.shape[1]), seeded_normal(40 + seed, X.shape[1]) * 2
        _, _, w, G, Gi, a = _closed_forms(X, q, p)
        b = X @ (Gi @ p)
        dTdq = 0.5 * X.T @ (w * (a - b * b))
        assert rel_err(grad(kinetic_energy, argnums=0)(q, p), dTdq) < 1e-11
        assert rel_err(grad(kinetic_energy, argnums=1)(q, p), Gi @ p) < 1e-12
        assert rel_err(inv_metric_vp(q, p), Gi @ p) < 1e-12
        h = 1e-6
        fd = torch.stack([(kinetic_energy(q + h * e, p) - kinetic_energy(q - h * e, p)) / (2 * h)
                          for e in torch.eye(X.shape[1], dtype=torch.float64)])
        assert rel_err(fd, dTdq) < 1e-7
        # T = 1/2 p^T G^-1 p + 1/2 log det G + (D/2) log 2 pi
        D = X.shape[1]
        T = 0.5 * p @ Gi @ p + 0.5 * torch.logdet(G) + 0.5 * D * np.log(2 * np.pi)
        assert abs(kinetic_energy(q, p).item() - T.item()) < 1e-11 * abs(T.item())


def test_lmc_christoffel_and_logdet_gradient_closed_forms():
    M, X, _ = _problem()
    _, kinetic, omega_tilde_fn, grad_logdet, mvp = olm.gaussian_riemannian(M.fisher_metric_fn)
    for seed in range(3):
        q, v = seeded_normal(50 + seed, X.shape[1]), seeded_normal(60 + seed, X.shape[1])
        eps = 0.07
        _, _, w, G, Gi, a = _closed_forms(X, q, v)
        omega = 0.5 * (X.T * (w * (X @ v))) @ X
        assert rel_err(omega_tilde_fn(q, v, eps), G + 0.5 * eps * omega) < 1e-12
        assert rel_err(grad_logdet(q), X.T @ (w * a)) < 1e-11
        assert rel_err(mvp(q, v), G @ v) < 1e-12
        K = -0.5 * torch.logdet(G) + 0.5 * v @ G @ v
        assert abs(kinetic(q, v).item() - K.item()) < 1e-11 * max(1.0, abs(K.item()))
