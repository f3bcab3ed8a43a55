"""Target densities and metric functions, single-chain, float64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Every function here maps ONE chain's position ``q`` (shape [D]) to a scalar or
matrix and is batched over chains by ``torch.func.vmap`` in the samplers, just
as the reference batches with ``jax.vmap``.
"""
import math

import torch
from torch.func import grad

from . import DTYPE

LOG2PI = math.log(2.0 * math.pi)


def norm_logpdf(x, loc, scale):
    """``jax.scipy.stats.norm.logpdf``: -(log(2*pi*scale^2) + ((x-loc)/scale)^2)/2."""
    scale = torch.as_tensor(scale, dtype=DTYPE)
    return -0.5 * (torch.log(2.0 * math.pi * scale * scale) + ((x - loc) / scale) ** 2)


class NealFunnel:
    """Neal's funnel exactly as examples/funnel/main.py:28-54.

    The log-scale coordinate is the LAST one, v = theta[D-1] ~ N(0, sigma^2),
    and x_i | v ~ N(0, e^v) for i < D-1 (main.py:34-39).
    """

    def __init__(self, D=2, mean=0.0, sigma=3.0):
        self.D = D
        self.mean = mean
        self.sigma = sigma

    def logp(self, theta):  # examples/funnel/main.py:34-39
        D = self.D
        v = theta[D - 1]
        return norm_logpdf(v, 0.0, self.sigma) + torch.sum(
            norm_logpdf(theta[: D - 1], 0.0, torch.exp(0.5 * v))
        )

    def inverse_jacobian(self, theta):  # examples/funnel/main.py:41-49
        D = self.D
        e = torch.exp(-0.5 * theta[-1])
        upper = torch.cat(
            [e * torch.eye(D - 1, dtype=DTYPE), (-0.5 * e * theta[: D - 1])[:, None]], dim=1
        )
        lowest = torch.cat([torch.zeros(D - 1, dtype=DTYPE), torch.tensor([1.0 / self.sigma], dtype=DTYPE)])
        return torch.cat([upper, lowest[None, :]], dim=0)

    def fisher_metric_fn(self, theta):  # examples/funnel/main.py:51-54
        J = self.inverse_jacobian(theta)
        metric = J.T @ J
        return 0.5 * (metric + metric.T)


def softplus(x):
    """log(1 + e^x) written as max(x, 0) + log1p(exp(-|x|)) (no threshold cut-off)."""
    return torch.clamp(x, min=0.0) + torch.log1p(torch.exp(-torch.abs(x)))


class LogisticRegression:
    """Bayesian logistic regression, prior beta ~ N(0, prior_var * I).

    The reference keeps the model user-supplied; this is the Bernoulli-logit
    likelihood sum_n [y_n eta_n - log(1 + e^{eta_n})], eta = X beta, plus the
    full Gaussian prior log-density including its normalising constant.
    """

    def __init__(self, X, y, prior_var=100.0):
        self.X = torch.as_tensor(X, dtype=DTYPE)
        self.y = torch.as_tensor(y, dtype=DTYPE)
        self.prior_var = float(prior_var)
        self.N, self.D = self.X.shape

    def logp(self, beta):
        eta = self.X @ beta
        loglik = torch.sum(self.y * eta - softplus(eta))
        logprior = -0.5 * torch.dot(beta, beta) / self.prior_var - 0.5 * self.D * math.log(
            2.0 * math.pi * self.prior_var
        )
        return loglik + logprior

    def fisher_metric_fn(self, beta):
        """G = X^T diag(s(1-s)) X + I/prior_var, s = sigmoid(X beta)."""
        s = torch.sigmoid(self.X @ beta)
        lam = s * (1.0 - s)
        return (self.X.T * lam) @ self.X + torch.eye(self.D, dtype=DTYPE) / self.prior_var


def monge_metric_fn(logdensity_fn, alpha2, D):
    """The Monge metric G(q) = I + alpha2 * grad l(q) grad l(q)^T as a metric_fn.

    This is the user-supplied ``metric_fn`` the RMHMC configurations pass to
    geomjax.rmhmc (geomjax/rmhmc/rmhmc.py:286-311).  grad l is taken by
    autodiff, so derivatives of G with respect to q go through the Hessian of
    l automatically, as they would under JAX.
    """
    g = grad(logdensity_fn)
    eye = torch.eye(D, dtype=DTYPE)

    def metric_fn(q):
        gq = g(q)
        return eye + alpha2 * torch.outer(gq, gq)

    return metric_fn


def identity_metric_fn(D):
    eye = torch.eye(D, dtype=DTYPE)

    def metric_fn(q):
        return eye + 0.0 * q.sum()

    return metric_fn
