"""Manifold MALA with a dense position-dependent metric (oracle, float64, batched).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/lmcmc/mmala.py:32-148 and geomjax/lmcmc/diffusions.py:27-94:
the Euler step of the Riemannian overdamped Langevin diffusion
    x' ~ N(mu(x), 2 eps G(x)^-1),  mu(x) = x + eps G(x)^-1 grad l(x) + eps Gamma(x),
    Gamma_i(x) = sum_j d(G^-1)_ij / d x_j      (jacfwd of inv(metric_fn), einsum "ijj"),
with the draw x' = mu + chol(2 eps sym(inv G)) z (jax.random.multivariate_normal's
Cholesky method, z the standard normals) and the asymmetric Metropolis test of
geomjax/mcmc/proposal.py:124-160.

The reference pairs the transition energies so that the proposal's weight is
    prev - new = [-l(x') - log q(x | x')] - [-l(x) - log q(x' | x)] = -log alpha,
the Metropolis-Hastings log ratio with its sign flipped (mmala.py:74-113 with
proposal.py:151-158).  ``reference_faithful=True`` (the default) reproduces
that; ``False`` uses +log alpha.
"""
import math
from typing import NamedTuple

import torch
from torch.func import jacfwd, vmap

from . import DTYPE
from .linalg import cholesky, solve
from .mcmc import Proposal, proposal_from_energy_diff, static_binomial_sampling


class DiffusionState(NamedTuple):  # geomjax/lmcmc/diffusions.py:27-32
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    metric: torch.Tensor
    Gamma: torch.Tensor


class MMALAInfo(NamedTuple):  # geomjax/lmcmc/mmala.py:32-50
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    proposal: DiffusionState
    is_divergent: torch.Tensor
    weight: torch.Tensor


def get_Gamma(position, metric_fn):
    """geomjax/lmcmc/diffusions.py:83-94, dense branch: d_invg[i, j, k] =
    d (G^-1)_ij / d x_k by forward-mode autodiff, Gamma_i = sum_j d_invg[i, j, j]."""
    d_invg = jacfwd(lambda th: torch.linalg.inv(metric_fn(th)))(position)
    return torch.einsum("ijj->i", d_invg)


def init(position, logdensity_fn, metric_fn):
    """geomjax/lmcmc/mmala.py:53-60 for a batch [C, D] of positions."""
    position = torch.as_tensor(position, dtype=DTYPE)
    logdensity, grad = vmap(torch.func.grad_and_value(logdensity_fn))(position)[::-1]
    metric = vmap(metric_fn)(position)
    Gamma = torch.stack([get_Gamma(q, metric_fn) for q in position])
    return DiffusionState(position, logdensity, grad, metric, Gamma)


def _mean(state, step_size):
    """mu(x) = x + eps solve(G, g) + eps Gamma (mmala.py:91-97, diffusions.py:55-61)."""
    return state.position + step_size * solve(state.metric, state.logdensity_grad) + step_size * state.Gamma


def _sym_inv(metric):
    inv = torch.linalg.inv(metric)
    return 0.5 * (inv + inv.T)


def transition_energy(state, new_position, step_size):
    """geomjax/lmcmc/mmala.py:74-113 (one chain): -l(state) - log N(new; mu(state),
    2 eps sym(inv G(state)))."""
    mean = _mean(state, step_size)
    cov = (2.0 * step_size) * _sym_inv(state.metric)
    L = cholesky(cov)
    r = new_position - mean
    w = torch.linalg.solve_triangular(L, r[:, None], upper=False)[:, 0]
    D = r.shape[0]
    logpdf = -0.5 * torch.dot(w, w) - torch.log(torch.diagonal(L)).sum() - 0.5 * D * math.log(2.0 * math.pi)
    return -state.logdensity - logpdf


def build_kernel(divergence_threshold=1000.0, reference_faithful=True):
    """geomjax/lmcmc/mmala.py:63-148.  ``noise`` = (z [C, D], u [C]): the
    normals of the proposal draw and the uniforms of the accept test."""

    def kernel(noise, state, logdensity_fn, step_size, metric_fn):
        z, u = noise
        C = state.position.shape[0]
        outs, infos = [], []
        for c in range(C):
            s = DiffusionState(*(t[c] for t in state))
            # the Euler step (diffusions.py:35-80, dense branch)
            mean = _mean(s, step_size)
            L = cholesky((2.0 * step_size) * _sym_inv(s.metric))
            x1 = mean + L @ z[c]
            g1 = torch.func.grad(logdensity_fn)(x1)
            s1 = DiffusionState(x1, logdensity_fn(x1), g1, metric_fn(x1), get_Gamma(x1, metric_fn))
            new_energy = transition_energy(s, x1, step_size)
            prev_energy = transition_energy(s1, s.position, step_size)
            if reference_faithful:
                new_prop = proposal_from_energy_diff(prev_energy, new_energy, s1)
            else:
                new_prop = proposal_from_energy_diff(new_energy, prev_energy, s1)
            is_div = bool(-new_prop.weight > divergence_threshold)
            cur_prop = Proposal(s, torch.tensor(0.0, dtype=DTYPE), torch.tensor(0.0, dtype=DTYPE),
                                torch.tensor(-float("inf"), dtype=DTYPE))
            chosen, acc, p = static_binomial_sampling(u[c], cur_prop, new_prop)
            outs.append(chosen.state)
            infos.append(MMALAInfo(p, acc, s1, is_div, new_prop.weight))
        new_state = DiffusionState(*(torch.stack([getattr(o, f) for o in outs]) for f in DiffusionState._fields))
        info = MMALAInfo(torch.stack([torch.as_tensor(i.acceptance_rate) for i in infos]),
                         torch.stack([torch.as_tensor(i.is_accepted) for i in infos]),
                         DiffusionState(*(torch.stack([getattr(i.proposal, f) for i in infos])
                                          for f in DiffusionState._fields)),
                         torch.tensor([i.is_divergent for i in infos]),
                         torch.stack([torch.as_tensor(i.weight) for i in infos]))
        return new_state, info

    return kernel
