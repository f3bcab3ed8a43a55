"""Shared MCMC machinery (oracle, batched over chains, float64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/mcmc/proposal.py:22-185 (static part),
geomjax/mcmc/trajectory.py:121-139 and the Euclidean HMC kernel
geomjax/mcmc/{hmc,integrators,metrics}.py used by the reference test
tests/test_samplers.py as the comparator for RMHMC with G = I.
"""
from typing import Any, NamedTuple

import torch
from torch.func import grad, grad_and_value, vmap


class Proposal(NamedTuple):  # geomjax/mcmc/proposal.py:22-40
    state: Any
    energy: torch.Tensor
    weight: torch.Tensor
    sum_log_p_accept: torch.Tensor


def proposal_generator(energy):
    """geomjax/mcmc/proposal.py:43-84.  ``energy`` maps a batched state to [C]."""

    def new(state):
        e = energy(state)
        return Proposal(state, e, torch.zeros_like(e), torch.full_like(e, -float("inf")))

    def update(initial_energy, state):
        return proposal_from_energy_diff(initial_energy, energy(state), state)

    return new, update


def proposal_from_energy_diff(initial_energy, new_energy, state):
    """geomjax/mcmc/proposal.py:87-121: NaN energy difference becomes -inf."""
    delta = initial_energy - new_energy
    delta = torch.where(torch.isnan(delta), torch.full_like(delta, -float("inf")), delta)
    weight = delta
    sum_log_p_accept = torch.minimum(delta, torch.zeros_like(delta))
    return Proposal(state, new_energy, weight, sum_log_p_accept)


def _select(mask, a, b):
    """Per-chain select over NamedTuple states (the batched ``lax.cond``)."""
    if isinstance(a, tuple):
        return type(a)(*[_select(mask, x, y) for x, y in zip(a, b)])
    if not torch.is_tensor(a):
        return a
    m = mask.reshape(mask.shape + (1,) * (a.dim() - mask.dim()))
    return torch.where(m, a, b)


def static_binomial_sampling(u, proposal, new_proposal):
    """geomjax/mcmc/proposal.py:168-185 with the key replaced by its uniform.

    ``jax.random.bernoulli(key, p)`` is ``uniform(key) < p``, so with the
    uniform ``u`` fed from the host the decision is ``u < p_accept``.
    """
    p_accept = torch.clamp(torch.exp(new_proposal.weight), max=1.0)
    do_accept = u < p_accept
    return _select(do_accept, new_proposal, proposal), do_accept, p_accept


def static_integration(integrator, direction=1):
    """geomjax/mcmc/trajectory.py:121-139 (fori_loop over the steps)."""

    def integrate(initial_state, step_size, num_integration_steps):
        state = initial_state
        for _ in range(num_integration_steps):
            state = integrator(state, direction * step_size)
        return state

    return integrate


# ---------------------------------------------------------------------------
# Euclidean HMC with a diagonal inverse mass: the comparator of the reference
# test (tests/test_samplers.py:27-32, 56).
# ---------------------------------------------------------------------------


class IntegratorState(NamedTuple):  # geomjax/mcmc/integrators.py
    position: torch.Tensor
    momentum: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor


class HMCState(NamedTuple):
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor


def velocity_verlet(logdensity_fn, inverse_mass_matrix):
    """geomjax/mcmc/integrators.py:45-107 (a1 = 0, b1 = 1/2, a2 = 1)."""
    gv = vmap(grad_and_value(logdensity_fn))

    def one_step(state, step_size):
        momentum = state.momentum + 0.5 * step_size * state.logdensity_grad
        position = state.position + step_size * (inverse_mass_matrix * momentum)
        g, logdensity = gv(position)
        momentum = momentum + 0.5 * step_size * g
        velocity = inverse_mass_matrix * momentum
        return IntegratorState(position, momentum, velocity, logdensity, g)

    return one_step


def hmc_init(position, logdensity_fn):
    g, l = vmap(grad_and_value(logdensity_fn))(position)
    return HMCState(position, l, g)


def hmc_kernel(noise, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps,
               divergence_threshold=1000.0):
    """geomjax/mcmc/hmc.py:118-178 with diagonal inverse mass and host-fed noise."""
    z, u = noise
    momentum = torch.sqrt(1.0 / inverse_mass_matrix) * z
    velocity = inverse_mass_matrix * momentum
    init = IntegratorState(state.position, momentum, velocity, state.logdensity, state.logdensity_grad)

    def energy(s):
        return -s.logdensity + 0.5 * torch.sum(inverse_mass_matrix * s.momentum * s.momentum, dim=-1)

    end = static_integration(velocity_verlet(logdensity_fn, inverse_mass_matrix))(
        init, step_size, num_integration_steps
    )
    end = end._replace(momentum=-end.momentum, velocity=-end.velocity)
    new, update = proposal_generator(energy)
    p0 = new(init)
    p1 = update(p0.energy, end)
    sampled, do_accept, p_accept = static_binomial_sampling(u, p0, p1)
    s = sampled.state
    return HMCState(s.position, s.logdensity, s.logdensity_grad), (do_accept, p_accept, p1)
