"""Lagrangian Monte Carlo with the Monge metric (oracle, float64, batched).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/lmcmonge/metrics.py:44-60,104-284,
geomjax/lmcmonge/integrators.py:26-194 and geomjax/lmcmonge/lmc.py:32-237,
472-565.  The metric is G = diag(m) + alpha2 g g^T with g = grad l, and
L = 1 + alpha2 g^T M^-1 g = det G / det M (the reference's
``determinant_metric``).

Reference quirks reproduced by default (``reference_faithful=True``):
  Q1  half_step_omega's Sherman-Morrison correction lacks the alpha2 factor
      (geomjax/lmcmonge/integrators.py:186-189 vs :178, :192).
  Q2  the momentum's rank-1 term is scaled by an extra L
      (geomjax/lmcmonge/lmc.py:202-204 with metrics.py:243-255).
With ``reference_faithful=False`` both are written as the exact algebra
(alpha2 in the correction; momentum = M v + alpha2 (v.g) g).
"""
from typing import NamedTuple

import torch
from torch.func import grad, grad_and_value, jvp, vmap

from . import DTYPE
from .linalg import cholesky
from .mcmc import Proposal, proposal_generator, static_binomial_sampling, static_integration


class RiemannianIntegratorState(NamedTuple):  # geomjax/lmcmonge/integrators.py:26-44
    alpha2: torch.Tensor
    position: torch.Tensor
    momentum: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad_norm: torch.Tensor
    dl_ig: torch.Tensor
    Hdl_ig: torch.Tensor
    ig_Hdl_ig: torch.Tensor
    logdensity_hvp_velocity_norm: torch.Tensor
    determinant_metric: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCState(NamedTuple):  # geomjax/lmcmonge/lmc.py:32-44
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCInfo(NamedTuple):  # geomjax/lmcmonge/lmc.py:63-98
    velocity: torch.Tensor
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    is_divergent: torch.Tensor
    energy: torch.Tensor
    proposal: Proposal
    num_integration_steps: int


def hvp(f, primals, tangents):  # geomjax/util.py:120-121
    return jvp(grad(f), (primals,), (tangents,))[1]


def _dot(a, b):
    return torch.sum(a * b, dim=-1)


def gaussian_riemannian(alpha2, diagonal_covariance, reference_faithful=True):
    """geomjax/lmcmonge/metrics.py:104-264 with set_inverse_mass 'fixed' (:44-60).

    Functions act on batched [C, D] / [C] tensors.
    """
    inverse_diagonal_mass = torch.as_tensor(diagonal_covariance, dtype=DTYPE)
    diagonal_mass = torch.reciprocal(inverse_diagonal_mass)
    if inverse_diagonal_mass.dim() != 1:
        raise ValueError(
            "The mass matrix has the wrong number of dimensions:"
            f" expected 1, got {inverse_diagonal_mass.dim()}."
        )

    def velocity_generator(z, position, alpha2, dl_ig):
        # :155-166 - dense Cholesky of G^-1 = diag(M^-1) - alpha2 dl_ig dl_ig^T
        def one(zc, dc):
            inverse_metric = torch.diag(inverse_diagonal_mass) - alpha2 * torch.outer(dc, dc)
            return cholesky(inverse_metric) @ zc

        return vmap(one)(z, dl_ig)

    def kinetic_energy(velocity, alpha2, logdensity_grad, determinant_metric):  # :168-186
        L = determinant_metric
        g_sqrt = torch.sqrt(diagonal_mass)
        v_g = g_sqrt * velocity
        return (
            -0.5 * (torch.log(L) + torch.sum(torch.log(diagonal_mass)))
            + 0.5 * _dot(v_g, v_g)
            + 0.5 * L * alpha2 * _dot(velocity, logdensity_grad) ** 2.0
        )

    def set_weighted_gradient(logdensity_grad):  # :188-191
        return inverse_diagonal_mass * logdensity_grad

    def normalizing_constant(alpha2, logdensity_grad):  # :193-197
        return 1 + alpha2 * _dot(inverse_diagonal_mass * logdensity_grad, logdensity_grad)

    def metric_vector_product(velocity, alpha2, logdensity_grad, determinant_metric):  # :243-255
        if reference_faithful:
            L = determinant_metric
            return velocity * diagonal_mass + (alpha2 * L * _dot(velocity, logdensity_grad))[:, None] * logdensity_grad
        # exact G v, with logdensity_grad the raw gradient g
        return velocity * diagonal_mass + (alpha2 * _dot(velocity, logdensity_grad))[:, None] * logdensity_grad

    return (velocity_generator, kinetic_energy, set_weighted_gradient, normalizing_constant,
            metric_vector_product)


def half_step_omega(alpha2, v, J, dl, Hv, L, dl_ig, Hdl_ig, ig_Hdl_ig, eps, reference_faithful=True):
    """geomjax/lmcmonge/integrators.py:158-194, batched."""
    d1 = 1.0 + 0.5 * eps * alpha2 * _dot(Hv, dl_ig)
    J = J - torch.log(torch.abs(d1))
    sL = torch.sqrt(L)[:, None]
    dphi = -dl + alpha2 * Hdl_ig / sL
    dphi_ig = -dl_ig + alpha2 * ig_Hdl_ig / sL
    v_temp = v - 0.5 * eps * sL * (dphi_ig - alpha2 * _dot(dphi, dl_ig)[:, None] * dl_ig)
    corr = 0.5 * eps * _dot(v_temp, Hv) / d1
    if not reference_faithful:
        corr = alpha2 * corr
    v_update = v_temp - corr[:, None] * dl_ig
    J = J + torch.log(torch.abs(1.0 - 0.5 * eps * alpha2 * _dot(Hdl_ig, v_update)))
    return v_update, J


def lan_integrator(logdensity_fn, set_weighted_gradient, normalizing_constant, metric_vector_product,
                   reference_faithful=True):
    """geomjax/lmcmonge/integrators.py:52-155."""
    vg = vmap(grad_and_value(logdensity_fn))
    hvp_fn = vmap(lambda x, v: hvp(logdensity_fn, x, v))

    def one_step(state, step_size):
        (alpha2, position, momentum, velocity, logdensity, logdensity_grad_norm, dl_ig, Hdl_ig,
         ig_Hdl_ig, hvn, determinant_metric, volume_adjustment) = state
        velocity, volume_adjustment = half_step_omega(
            alpha2, velocity, volume_adjustment, logdensity_grad_norm, hvn, determinant_metric,
            dl_ig, Hdl_ig, ig_Hdl_ig, step_size, reference_faithful)
        position = position + step_size * velocity
        logdensity_grad, logdensity = vg(position)
        determinant_metric = normalizing_constant(alpha2, logdensity_grad)
        sdm = torch.sqrt(determinant_metric)[:, None]
        logdensity_grad_norm = logdensity_grad / sdm
        hvn = hvp_fn(position, velocity) / sdm
        dl_ig = set_weighted_gradient(logdensity_grad_norm)
        Hdl_ig = hvp_fn(position, dl_ig) / sdm
        ig_Hdl_ig = set_weighted_gradient(Hdl_ig)
        velocity, volume_adjustment = half_step_omega(
            alpha2, velocity, volume_adjustment, logdensity_grad_norm, hvn, determinant_metric,
            dl_ig, Hdl_ig, ig_Hdl_ig, step_size, reference_faithful)
        hvn = hvp_fn(position, velocity) / sdm
        momentum = metric_vector_product(velocity, alpha2, logdensity_grad, determinant_metric)
        return RiemannianIntegratorState(alpha2, position, momentum, velocity, logdensity,
                                         logdensity_grad_norm, dl_ig, Hdl_ig, ig_Hdl_ig, hvn,
                                         determinant_metric, volume_adjustment)

    return one_step


def lmcmonge_energy(kinetic_energy):  # geomjax/lmcmonge/metrics.py:267-284
    def energy(state):
        return (-state.logdensity
                + kinetic_energy(state.velocity, state.alpha2, state.logdensity_grad_norm,
                                 state.determinant_metric)
                - state.volume_adjustment)

    return energy


def init(position, logdensity_fn):  # geomjax/lmcmonge/lmc.py:101-109
    g, l = vmap(grad_and_value(logdensity_fn))(position)
    return LMCState(position, l, g, torch.zeros_like(l))


def flip_velocity(state):  # geomjax/lmcmonge/lmc.py:539-565
    return state._replace(momentum=-1.0 * state.momentum, velocity=-1.0 * state.velocity)


def lmc_proposal(integrator, kinetic_energy, step_size, num_integration_steps=1,
                 divergence_threshold=1000.0):  # geomjax/lmcmonge/lmc.py:472-536
    build_trajectory = static_integration(integrator)
    init_proposal, generate_proposal = proposal_generator(lmcmonge_energy(kinetic_energy))

    def generate(u, state):
        end_state = build_trajectory(state, step_size, num_integration_steps)
        end_state = flip_velocity(end_state)
        proposal = init_proposal(state)
        new_proposal = generate_proposal(proposal.energy, end_state)
        is_diverging = -new_proposal.weight > divergence_threshold
        sampled, do_accept, p_accept = static_binomial_sampling(u, proposal, new_proposal)
        info = LMCInfo(state.velocity, p_accept, do_accept, is_diverging, new_proposal.energy,
                       new_proposal, num_integration_steps)
        return sampled.state, info

    return generate


def build_kernel(divergence_threshold=1000.0, reference_faithful=True):
    """geomjax/lmcmonge/lmc.py:130-237 with the key replaced by noise=(z, u)."""

    def kernel(noise, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps,
               alpha2):
        z, u = noise
        position, logdensity, logdensity_grad, volume_adjustment = state
        (velocity_generator, kinetic_energy, set_weighted_gradient, normalizing_constant,
         metric_vector_product) = gaussian_riemannian(alpha2, inverse_mass_matrix, reference_faithful)
        determinant_metric = normalizing_constant(alpha2, logdensity_grad)
        sdm = torch.sqrt(determinant_metric)[:, None]
        logdensity_grad_norm = logdensity_grad / sdm
        integrator = lan_integrator(logdensity_fn, set_weighted_gradient, normalizing_constant,
                                    metric_vector_product, reference_faithful)
        generate = lmc_proposal(integrator, kinetic_energy, step_size, num_integration_steps,
                                divergence_threshold)
        hvp_fn = vmap(lambda x, v: hvp(logdensity_fn, x, v))
        dl_ig = set_weighted_gradient(logdensity_grad_norm)
        Hdl_ig = hvp_fn(position, dl_ig) / sdm
        ig_Hdl_ig = set_weighted_gradient(Hdl_ig)
        velocity = velocity_generator(z, position, alpha2, dl_ig)
        momentum = metric_vector_product(velocity, alpha2, logdensity_grad, determinant_metric)
        hvn = hvp_fn(position, velocity) / sdm
        a2 = torch.as_tensor(alpha2, dtype=DTYPE)
        istate = RiemannianIntegratorState(a2, position, momentum, velocity, logdensity,
                                           logdensity_grad_norm, dl_ig, Hdl_ig, ig_Hdl_ig, hvn,
                                           determinant_metric, volume_adjustment)
        proposal, info = generate(u, istate)
        sdm1 = torch.sqrt(proposal.determinant_metric)[:, None]
        new_state = LMCState(proposal.position, proposal.logdensity,
                             proposal.logdensity_grad_norm * sdm1, proposal.volume_adjustment)
        return new_state, info

    return kernel
