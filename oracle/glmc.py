"""Generalized LMC with a persistent velocity and a non-reversible slice (oracle,
float64, batched).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/lmcmc/glmc.py:33-186 and the non-reversible slice sampling of
geomjax/mcmc/proposal.py:262-277 on top of the lmcmc oracle (the same metric
functions, lan_integrator, energy and flip).  One correction: the reference
builds the six-field IntegratorState of geomjax/lmcmc/integrators.py:26-38 from
five values (glmc.py:170-172, a TypeError at the first transition); here the
state carries the momentum G(q) v, built as the lmcmc kernel builds it
(geomjax/lmcmc/lmc.py:170-174).  The key is replaced by its draws: the
standard normals z of the velocity refresh and the value of the slice noise.
"""
from typing import NamedTuple

import torch
from torch.func import grad_and_value, vmap

from .lmcmc import LMCInfo, IntegratorState, flip_velocity, gaussian_riemannian, lan_integrator, lmc_energy
from .mcmc import _select, proposal_generator, static_integration


class GLMCState(NamedTuple):  # geomjax/lmcmc/glmc.py:33-55
    position: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    slice: torch.Tensor
    volume_adjustment: torch.Tensor


def init(position, z, u, logdensity_fn):
    """geomjax/lmcmc/glmc.py:58-71: the velocity is a standard normal draw z (not
    scaled by the metric) and the slice uniform on [-1, 1), -1 + 2u."""
    g, l = vmap(grad_and_value(logdensity_fn))(position)
    return GLMCState(position, z, l, g, -1.0 + 2.0 * u, torch.zeros_like(l))


def nonreversible_slice_sampling(slice, proposal, new_proposal):  # geomjax/mcmc/proposal.py:262-277
    delta_energy = new_proposal.weight
    do_accept = torch.log(torch.abs(slice)) <= delta_energy
    new_slice = torch.where(do_accept, slice * torch.exp(-delta_energy), slice)
    return _select(do_accept, new_proposal, proposal), do_accept, new_slice


def update_velocity(z, position, velocity, alpha, velocity_generator):  # geomjax/lmcmc/glmc.py:172-190
    fresh = vmap(velocity_generator)(z, position)
    return velocity * torch.sqrt(torch.tensor(1.0 - alpha, dtype=velocity.dtype)) + \
        torch.sqrt(torch.tensor(alpha, dtype=velocity.dtype)) * fresh


def build_kernel(divergence_threshold=1000.0):
    """geomjax/lmcmc/glmc.py:74-169 with ``noise`` = (z, slice_noise)."""

    def kernel(noise, state, logdensity_fn, step_size, metric_fn, alpha, delta):
        z, slice_noise = noise
        (velocity_generator, kinetic_energy, omega_tilde_fn, grad_logdetmetric,
         metric_vector_product) = gaussian_riemannian(metric_fn)
        integrator = lan_integrator(logdensity_fn, omega_tilde_fn, grad_logdetmetric, metric_vector_product)
        position, velocity, logdensity, logdensity_grad, slice, volume_adjustment = state
        velocity = update_velocity(z, position, velocity, alpha, velocity_generator)
        slice = torch.remainder(slice + 1.0 + delta + slice_noise, 2.0) - 1.0  # jnp's % takes the divisor's sign
        momentum = vmap(metric_vector_product)(position, velocity)
        istate = IntegratorState(position, momentum, velocity, logdensity, logdensity_grad, volume_adjustment)
        # lmc_proposal (geomjax/lmcmc/lmc.py:446-475) with one step and the slice sampler
        end = flip_velocity(static_integration(integrator)(istate, step_size, 1))
        new, update = proposal_generator(lmc_energy(kinetic_energy))
        proposal = new(istate)
        new_proposal = update(proposal.energy, end)
        is_diverging = -new_proposal.weight > divergence_threshold
        sampled, do_accept, p_accept = nonreversible_slice_sampling(slice, proposal, new_proposal)
        info = LMCInfo(istate.velocity, p_accept, do_accept, is_diverging, new_proposal.energy, new_proposal, 1)
        st = flip_velocity(sampled.state)
        return GLMCState(st.position, st.velocity, st.logdensity, st.logdensity_grad, info.acceptance_rate,
                         st.volume_adjustment), info

    return kernel
