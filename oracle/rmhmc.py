"""RMHMC with the implicit midpoint integrator (oracle, float64, batched).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/rmhmc/metrics.py:42-129, geomjax/rmhmc/integrators.py:25-156
and geomjax/rmhmc/rmhmc.py:30-176,376-462.  Single-chain functions are batched
with ``torch.func.vmap``; the fixed-point ``while_loop`` is written at the batch
level with the per-chain freeze that ``jax.vmap`` gives a batched while_loop:
the body runs while any chain's condition holds and only chains whose
condition holds take the new carry.
"""
import math
from typing import Any, NamedTuple

import torch
from torch.func import grad, grad_and_value, vmap

from . import DTYPE
from .linalg import cholesky, solve
from .mcmc import Proposal, proposal_generator, static_binomial_sampling, static_integration

LOG2PI = math.log(2.0 * math.pi)


class IntegratorState(NamedTuple):  # geomjax/rmhmc/integrators.py:25-36
    position: torch.Tensor
    momentum: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor


class FixedPointIterationInfo(NamedTuple):  # geomjax/rmhmc/integrators.py:47-50
    success: torch.Tensor
    norm: torch.Tensor
    iters: torch.Tensor


class RMHMCState(NamedTuple):  # geomjax/rmhmc/rmhmc.py:30-41
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor


class RMHMCInfo(NamedTuple):  # geomjax/rmhmc/rmhmc.py:58-93
    momentum: torch.Tensor
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    is_divergent: torch.Tensor
    energy: torch.Tensor
    proposal: Proposal
    num_integration_steps: int


def mvn_logpdf(x, cov):
    """``jax.scipy.stats.multivariate_normal.logpdf(x, 0, cov)`` for one vector.

    Cholesky of ``cov``, a lower triangular solve, then
    -|y|^2/2 - (n/2) log(2 pi) - sum log diag(L).
    """
    n = x.shape[-1]
    L = cholesky(cov)
    y = torch.linalg.solve_triangular(L, x[:, None], upper=False)[:, 0]
    return -0.5 * torch.dot(y, y) - 0.5 * n * LOG2PI - torch.sum(torch.log(torch.diagonal(L)))


def gaussian_riemannian(metric_fn):
    """geomjax/rmhmc/metrics.py:42-129 (single chain).

    ``momentum_generator`` takes the standard-normal draw ``z`` in place of the
    key (geomjax/util.py:58-83: mu + sigma @ z with sigma = chol(G)).
    """

    def momentum_generator(z, position):
        mass_matrix = metric_fn(position)
        if mass_matrix.dim() == 1:
            return torch.sqrt(mass_matrix) * z
        if mass_matrix.dim() == 2:
            return cholesky(mass_matrix) @ z
        raise ValueError(
            "The mass matrix has the wrong number of dimensions:"
            f" expected 1 or 2, got {mass_matrix.dim()}."
        )

    def kinetic_energy(position, momentum):
        mass_matrix = metric_fn(position)
        if mass_matrix.dim() == 1:
            return -torch.sum(
                -0.5 * (torch.log(2.0 * math.pi * mass_matrix) + momentum * momentum / mass_matrix)
            )
        if mass_matrix.dim() == 2:
            return -mvn_logpdf(momentum, mass_matrix)
        raise ValueError(
            "The mass matrix has the wrong number of dimensions:"
            f" expected 1 or 2, got {mass_matrix.dim()}."
        )

    def inverse_metric_vector_product(position, momentum):
        metric = metric_fn(position)
        if metric.dim() == 1:
            return momentum / metric
        return solve(metric_fn(position), momentum)

    return momentum_generator, kinetic_energy, inverse_metric_vector_product


def _max_abs_diff(x, xp):
    """norm_fn(ravel(x - xp)) = max |.| over the concatenation of q and p."""
    return torch.maximum(
        torch.amax(torch.abs(x[0] - xp[0]), dim=-1), torch.amax(torch.abs(x[1] - xp[1]), dim=-1)
    )


def solve_fixed_point_iteration(func, x0, *, convergence_tol=1e-6, divergence_tol=1e10, max_iters=100):
    """geomjax/rmhmc/integrators.py:53-89, batched over chains.

    First evaluation outside the loop (:84), initial norm |x1 - x0| (:86),
    loop while n < max_iters & finite & norm < divergence_tol &
    norm > convergence_tol (:67-74), success = finite & norm <= tol (:88).
    """
    x, aux = func(x0)
    norm = _max_abs_diff(x, x0)
    n = torch.zeros(norm.shape, dtype=torch.int64)
    while True:
        active = (n < max_iters) & torch.isfinite(norm) & (norm < divergence_tol) & (norm > convergence_tol)
        if not bool(active.any()):
            break
        xn, auxn = func(x)
        normn = _max_abs_diff(xn, x)
        m = active[:, None]
        x = (torch.where(m, xn[0], x[0]), torch.where(m, xn[1], x[1]))
        aux = torch.where(m, auxn, aux)
        norm = torch.where(active, normn, norm)
        n = n + active.to(torch.int64)
    success = torch.isfinite(norm) & (norm <= convergence_tol)
    return x, aux, FixedPointIterationInfo(success, norm, n)


def implicit_midpoint(logdensity_fn, kinetic_energy_fn, inverse_metric_vector_product, *,
                      solver=solve_fixed_point_iteration, return_info=False, **solver_kwargs):
    """geomjax/rmhmc/integrators.py:92-156 (Brofos & Lederman implicit midpoint).

    x^{n+1} = z0 + (eps/2) F(x^n), F = (dT/dp, grad l - dT/dq), solved by
    fixed-point iteration; then ONE explicit half step from the last iterate
    x* (:146-148), velocity G(q1)^-1 p1 (:150), and (l, grad l) at q1 (:154).
    """
    grad_logp = vmap(grad(logdensity_fn))
    value_and_grad_logp = vmap(grad_and_value(logdensity_fn))
    kinetic_energy_grad_fn = vmap(grad(lambda q, p: kinetic_energy_fn(q, p), argnums=(0, 1)))
    imvp = vmap(inverse_metric_vector_product)

    def one_step(state, step_size):
        position, momentum = state.position, state.momentum

        def _update(q, p, dUdq, initial):
            dTdq, dHdp = kinetic_energy_grad_fn(q, p)
            dHdq = dTdq - dUdq
            return initial[0] + 0.5 * step_size * dHdp, initial[1] - 0.5 * step_size * dHdq

        def _step(args):
            q, p = args
            dLdq = grad_logp(q)
            return _update(q, p, dLdq, (position, momentum)), dLdq

        (q, p), dLdq, info = solver(_step, (position, momentum), **solver_kwargs)
        dLdq = grad_logp(q)
        q, p = _update(q, p, dLdq, (q, p))
        v = imvp(q, p)
        g, l = value_and_grad_logp(q)
        new_state = IntegratorState(q, p, v, l, g)
        if return_info:
            return new_state, info
        return new_state

    return one_step


def hmc_energy(kinetic_energy):
    """geomjax/mcmc/metrics.py:160-166: H = -l + T(q, p)."""
    ke = vmap(kinetic_energy)

    def energy(state):
        return -state.logdensity + ke(state.position, state.momentum)

    return energy


def init(position, logdensity_fn):  # geomjax/rmhmc/rmhmc.py:96-98
    g, l = vmap(grad_and_value(logdensity_fn))(position)
    return RMHMCState(position, l, g)


def flip_momentum(state):  # geomjax/rmhmc/rmhmc.py:443-462
    return IntegratorState(state.position, -1.0 * state.momentum, -1.0 * state.velocity,
                           state.logdensity, state.logdensity_grad)


def rmhmc_proposal(integrator, kinetic_energy, step_size, num_integration_steps=1,
                   divergence_threshold=1000.0, *, trace=None):
    """geomjax/rmhmc/rmhmc.py:376-440.

    ``trace`` (a list) collects the per-step FixedPointIterationInfo when the
    integrator was built with return_info=True.
    """
    init_proposal, generate_proposal = proposal_generator(hmc_energy(kinetic_energy))

    def traced(state, eps):
        out = integrator(state, eps)
        if isinstance(out, tuple) and len(out) == 2 and isinstance(out[1], FixedPointIterationInfo):
            if trace is not None:
                trace.append(out[1])
            return out[0]
        return out

    build_trajectory = static_integration(traced)

    def generate(u, state):
        end_state = build_trajectory(state, step_size, num_integration_steps)
        end_state = flip_momentum(end_state)
        proposal = init_proposal(state)
        new_proposal = generate_proposal(proposal.energy, end_state)
        is_diverging = -new_proposal.weight > divergence_threshold
        sampled, do_accept, p_accept = static_binomial_sampling(u, proposal, new_proposal)
        info = RMHMCInfo(state.momentum, p_accept, do_accept, is_diverging, new_proposal.energy,
                         new_proposal, num_integration_steps)
        return sampled.state, info

    return generate


def build_kernel(integrator=implicit_midpoint, divergence_threshold=1000.0, **solver_kwargs):
    """geomjax/rmhmc/rmhmc.py:110-176 with the key replaced by noise=(z, u).

    ``solver_kwargs`` (max_iters, convergence_tol, divergence_tol) are bound
    into the integrator the way a caller binds them with functools.partial in
    the reference.  Pass ``trace=[]`` to collect fixed-point infos.
    """

    def kernel(noise, state, logdensity_fn, step_size, metric_fn, num_integration_steps, trace=None):
        z, u = noise
        momentum_generator, kinetic_energy_fn, imvp = gaussian_riemannian(metric_fn)
        symplectic_integrator = integrator(
            logdensity_fn, kinetic_energy_fn, imvp, return_info=trace is not None, **solver_kwargs
        )
        generate = rmhmc_proposal(symplectic_integrator, kinetic_energy_fn, step_size,
                                  num_integration_steps, divergence_threshold, trace=trace)
        position, logdensity, logdensity_grad = state
        momentum = vmap(momentum_generator)(z, position)
        velocity = vmap(imvp)(position, momentum)
        integrator_state = IntegratorState(position, momentum, velocity, logdensity, logdensity_grad)
        proposal, info = generate(u, integrator_state)
        return RMHMCState(proposal.position, proposal.logdensity, proposal.logdensity_grad), info

    return kernel


class rmhmc:
    """geomjax/rmhmc/rmhmc.py:247-311 (init_fn / step_fn pair)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, num_integration_steps, *,
                divergence_threshold=1000.0, **solver_kwargs):
        kernel = cls.build_kernel(divergence_threshold=divergence_threshold, **solver_kwargs)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(noise, state, trace=None):
            return kernel(noise, state, logdensity_fn, step_size, metric_fn, num_integration_steps,
                          trace=trace)

        return init_fn, step_fn
