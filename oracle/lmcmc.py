"""Lagrangian Monte Carlo with a general dense metric (oracle, float64, batched).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/lmcmc/metrics.py:42-221, geomjax/lmcmc/integrators.py:26-144
and geomjax/lmcmc/lmc.py:30-182,411-499.  The Christoffel contraction
Omega~(v) is built from the full D x D x D ``jacfwd`` of the metric, and the
volume terms from LU factorisations, exactly as the reference does.
"""
from typing import NamedTuple

import torch
from torch.func import grad, grad_and_value, jacfwd, vmap

from . import DTYPE
from .linalg import cholesky, lu_factor
from .mcmc import Proposal, proposal_generator, static_binomial_sampling, static_integration


class IntegratorState(NamedTuple):  # geomjax/lmcmc/integrators.py:26-38
    position: torch.Tensor
    momentum: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCState(NamedTuple):  # geomjax/lmcmc/lmc.py:30-42
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCInfo(NamedTuple):  # geomjax/lmcmc/lmc.py:60-95
    velocity: torch.Tensor
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    is_divergent: torch.Tensor
    energy: torch.Tensor
    proposal: Proposal
    num_integration_steps: int


def gaussian_riemannian(metric_fn):
    """geomjax/lmcmc/metrics.py:42-206, single-chain functions."""

    def velocity_generator(z, position):  # :75-91
        metric = metric_fn(position)
        metric = 0.5 * (metric + metric.T)
        L = cholesky(metric)
        eye = torch.eye(metric.shape[0], dtype=DTYPE)
        metric_invsqrt = torch.linalg.solve_triangular(L.T, eye, upper=True)  # L^-T
        return metric_invsqrt @ z

    def kinetic_energy(position, velocity):  # :93-112
        metric = metric_fn(position)
        metric = 0.5 * (metric + metric.T)
        _, logdetG = torch.linalg.slogdet(metric)
        return -0.5 * logdetG + 0.5 * torch.dot(metric @ velocity, velocity)

    def omega_tilde_fn(position, velocity, step_size):  # :158-179
        metric = metric_fn(position)
        d_g = jacfwd(metric_fn)(position)  # d_g[a, b, i] = dG_ab / dq_i
        partial_1 = torch.einsum("i,jli->lj", velocity, d_g)
        partial_2 = torch.einsum("i,ilj->lj", velocity, d_g)
        partial_3 = torch.einsum("i,ijl->lj", velocity, d_g)
        omega_tilde = 0.5 * (partial_1 + partial_2 - partial_3)
        return metric + 0.5 * step_size * omega_tilde

    def grad_logdetmetric(position):  # :181-189
        return grad(lambda theta: torch.linalg.slogdet(metric_fn(theta))[1])(position)

    def metric_vector_product(position, velocity):  # :191-197
        return metric_fn(position) @ velocity

    return velocity_generator, kinetic_energy, omega_tilde_fn, grad_logdetmetric, metric_vector_product


def _logabsdet_lu(A):
    LU, _piv = lu_factor(A)
    return LU, _piv, torch.sum(torch.log(torch.abs(torch.diagonal(LU))))


def lan_integrator(logdensity_fn, omega_tilde_fn, grad_logdetmetric, metric_vector_product):
    """geomjax/lmcmc/integrators.py:51-144 (single-chain half step, vmapped)."""
    vg = vmap(grad_and_value(logdensity_fn))
    mvp = vmap(metric_vector_product)

    def half_step_fn(position, velocity, volume_adjustment, logdensity_grad, step_size):  # :61-91
        A = omega_tilde_fn(position, velocity, step_size)
        LU, piv, logdet = _logabsdet_lu(A)
        volume_adjustment = volume_adjustment - logdet
        dphi = -logdensity_grad + 0.5 * grad_logdetmetric(position)
        v_temp = metric_vector_product(position, velocity) - 0.5 * step_size * dphi
        v_update = torch.linalg.lu_solve(LU, piv, v_temp[:, None])[:, 0]
        A = omega_tilde_fn(position, v_update, -step_size)
        _, _, logdet = _logabsdet_lu(A)
        volume_adjustment = volume_adjustment + logdet
        return v_update, volume_adjustment

    half = vmap(half_step_fn, in_dims=(0, 0, 0, 0, None))

    def one_step(state, step_size):  # :93-142
        position, _, velocity, logdensity, logdensity_grad, volume_adjustment = state
        velocity, volume_adjustment = half(position, velocity, volume_adjustment, logdensity_grad, step_size)
        position = position + step_size * velocity
        logdensity_grad, logdensity = vg(position)
        velocity, volume_adjustment = half(position, velocity, volume_adjustment, logdensity_grad, step_size)
        momentum = mvp(position, velocity)
        return IntegratorState(position, momentum, velocity, logdensity, logdensity_grad, volume_adjustment)

    return one_step


def lmc_energy(kinetic_energy):  # geomjax/lmcmc/metrics.py:209-221
    ke = vmap(kinetic_energy)

    def energy(state):
        return -state.logdensity + ke(state.position, state.velocity) - state.volume_adjustment

    return energy


def init(position, logdensity_fn):  # geomjax/lmcmc/lmc.py:98-101
    g, l = vmap(grad_and_value(logdensity_fn))(position)
    return LMCState(position, l, g, torch.zeros_like(l))


def flip_velocity(state):  # geomjax/lmcmc/lmc.py:478-499
    return state._replace(momentum=-1.0 * state.momentum, velocity=-1.0 * state.velocity)


def lmc_proposal(integrator, kinetic_energy, step_size, num_integration_steps=1,
                 divergence_threshold=1000.0):  # geomjax/lmcmc/lmc.py:411-475
    build_trajectory = static_integration(integrator)
    init_proposal, generate_proposal = proposal_generator(lmc_energy(kinetic_energy))

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


def build_kernel(divergence_threshold=1000.0):
    """geomjax/lmcmc/lmc.py:114-182 with the key replaced by noise=(z, u)."""

    def kernel(noise, state, logdensity_fn, step_size, metric_fn, num_integration_steps):
        z, u = noise
        (velocity_generator, kinetic_energy, omega_tilde_fn, grad_logdetmetric,
         metric_vector_product) = gaussian_riemannian(metric_fn)
        integrator = lan_integrator(logdensity_fn, omega_tilde_fn, grad_logdetmetric, metric_vector_product)
        generate = lmc_proposal(integrator, kinetic_energy, step_size, num_integration_steps,
                                divergence_threshold)
        position, logdensity, logdensity_grad, volume_adjustment = state
        velocity = vmap(velocity_generator)(z, position)
        momentum = vmap(metric_vector_product)(position, velocity)
        istate = IntegratorState(position, momentum, velocity, logdensity, logdensity_grad, volume_adjustment)
        proposal, info = generate(u, istate)
        return LMCState(proposal.position, proposal.logdensity, proposal.logdensity_grad,
                        proposal.volume_adjustment), info

    return kernel
