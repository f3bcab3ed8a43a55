"""NUTS with the RMHMC implicit-midpoint integrator (oracle, float64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates geomjax/rmhmc/nuts.py:78-161,253-348 (kernel, iterative proposal),
geomjax/lmcmonge/nuts.py:80-207,300-404 (the LMC-Monge kernel),
geomjax/lmcmc/nuts.py:78-161,253-351 (the general-metric LMC kernel),
geomjax/mcmc/trajectory.py:40-103 (trajectory helpers), :149-317
(dynamic_progressive_integration), :517-668 (dynamic_multiplicative_expansion),
geomjax/mcmc/termination.py:30-119 (numpyro-style iterative U-turn),
geomjax/mcmc/proposal.py:43-121,198-255 (proposals, progressive uniform and
biased sampling) and the U-turn criteria of geomjax/rmhmc/metrics.py:100-117.

Chains run one at a time (each has its own trajectory); the integrator is the
batched implicit midpoint of oracle.rmhmc on a batch of one.  The JAX keys
are replaced by explicit uniforms, consumed in this order per chain:
``u_dir[e]`` (direction of expansion e: +1 when < 0.5), ``u_merge[e]`` (the
biased merge of expansion e), and ``u_step[j]`` for the j-th integration step
of the transition (progressive uniform sampling inside a subtree; the first
step of every subtree consumes its uniform without using it).
"""
import math
from typing import Any, NamedTuple

import torch
from torch.func import vmap

from .mcmc import Proposal
from .rmhmc import IntegratorState, RMHMCState, gaussian_riemannian, hmc_energy, implicit_midpoint


class NUTSInfo(NamedTuple):  # geomjax/rmhmc/nuts.py:36-75
    momentum: Any
    is_divergent: Any
    is_turning: Any
    energy: Any
    trajectory_leftmost_state: Any
    trajectory_rightmost_state: Any
    num_trajectory_expansions: Any
    num_integration_steps: Any
    acceptance_rate: Any


class Trajectory(NamedTuple):  # geomjax/mcmc/trajectory.py:40-45
    leftmost_state: IntegratorState
    rightmost_state: IntegratorState
    momentum_sum: torch.Tensor
    velocity_sum: torch.Tensor
    num_states: int


def _is_turning(criterion, m_left, m_right, m_sum, v_left, v_right, v_sum):
    """geomjax/rmhmc/metrics.py:100-117: "euc" rho = velocity sum against the end
    velocities, "bet" rho = momentum sum against the end velocities, "riem"
    rho = velocity sum against the end momenta."""
    if criterion == "euc":
        rho, left, right = v_sum, v_left, v_right
    elif criterion == "bet":
        rho, left, right = m_sum, v_left, v_right
    elif criterion == "riem":
        rho, left, right = v_sum, m_left, m_right
    else:
        raise ValueError(criterion)
    return bool(torch.dot(left, rho) <= 0) or bool(torch.dot(right, rho) <= 0)


def _leaf_idx_to_ckpt_idxs(n):
    """geomjax/mcmc/termination.py:71-87."""
    idx_max = bin(n >> 1).count("1")
    num_subtrees = 0
    m = n
    while m & 1:
        num_subtrees += 1
        m >>= 1
    return idx_max - num_subtrees + 1, idx_max


def _proposal(H0, energy, state):
    """geomjax/mcmc/proposal.py:87-121: weight = H0 - H (NaN -> -inf), sum_log_p_accept = min(weight, 0)."""
    e = float(energy(state)[0])
    w = H0 - e
    if math.isnan(w):
        w = -math.inf
    return Proposal(state, e, w, min(w, 0.0))


def _logaddexp(a, b):
    return float(torch.logaddexp(torch.tensor(a, dtype=torch.float64), torch.tensor(b, dtype=torch.float64)))


def _sample(u, proposal, new, biased):
    """progressive_uniform_sampling (:198-222) / progressive_biased_sampling (:225-254)."""
    dw = new.weight - proposal.weight
    if biased:
        p = min(1.0, math.exp(dw)) if dw < 700 else 1.0
    else:
        p = 1.0 / (1.0 + math.exp(-dw)) if dw > -700 else 0.0
    w = _logaddexp(proposal.weight, new.weight)
    s = _logaddexp(proposal.sum_log_p_accept, new.sum_log_p_accept)
    if u < p:
        return Proposal(new.state, new.energy, w, s)
    return Proposal(proposal.state, proposal.energy, w, s)


def _expand(s0, H0, one_step, energy, step_size, u_dir, u_merge, u_step, M, criterion, divergence_threshold):
    """One chain's iterative NUTS expansion from s0 (a batch-of-one integrator state
    with ``momentum`` and ``velocity``): geomjax/mcmc/trajectory.py:149-317,517-668
    with the numpyro U-turn check of geomjax/mcmc/termination.py:30-119.
    Returns (proposal, trajectory, expansions, diverging, turning)."""
    D = s0.position.shape[-1]
    dtype = s0.position.dtype
    proposal = Proposal(s0, H0, 0.0, -math.inf)
    traj = Trajectory(s0, s0, s0.momentum[0].clone(), s0.velocity[0].clone(), 0)
    r_ck = torch.zeros(M, D, dtype=dtype)
    rs_ck = torch.zeros(M, D, dtype=dtype)
    v_ck = torch.zeros(M, D, dtype=dtype)
    vs_ck = torch.zeros(M, D, dtype=dtype)
    j = 0  # integration steps taken in this transition (index into u_step)
    e = 0
    diverging = turning = False
    while e < M and not diverging and not turning:
        direction = 1 if float(u_dir[e]) < 0.5 else -1
        start = traj.rightmost_state if direction > 0 else traj.leftmost_state
        # ---- dynamic_progressive_integration (:149-317) ----
        max_steps = 2 ** e
        sub_prop = _proposal(H0, energy, start)
        sub = Trajectory(start, start, start.momentum[0].clone(), start.velocity[0].clone(), 0)
        step, div_s, term_s = 0, False, False
        while step < max_steps and not term_s and not div_s:
            new = one_step(sub.rightmost_state, direction * step_size)
            new_prop = _proposal(H0, energy, new)
            div_s = -new_prop.weight > divergence_threshold
            if step == 0:
                sub = Trajectory(new, new, new.momentum[0].clone(), new.velocity[0].clone(), 1)
                sub_prop = new_prop
            else:
                sub = Trajectory(sub.leftmost_state, new, sub.momentum_sum + new.momentum[0],
                                 sub.velocity_sum + new.velocity[0], sub.num_states + 1)
                sub_prop = _sample(float(u_step[j]), sub_prop, new_prop, biased=False)
            j += 1
            # iterative U-turn (termination.py:30-69, 89-117)
            idx = _leaf_idx_to_ckpt_idxs(step)
            if step % 2 == 0:
                r_ck[idx[1]] = new.momentum[0]
                rs_ck[idx[1]] = sub.momentum_sum
                v_ck[idx[1]] = new.velocity[0]
                vs_ck[idx[1]] = sub.velocity_sum
            i = idx[1]
            while i >= idx[0] and not term_s:
                term_s = _is_turning(criterion, r_ck[i], new.momentum[0], sub.momentum_sum - rs_ck[i] + r_ck[i],
                                     v_ck[i], new.velocity[0], sub.velocity_sum - vs_ck[i] + v_ck[i])
                i -= 1
            step += 1
        if direction < 0:
            sub = Trajectory(sub.rightmost_state, sub.leftmost_state, sub.momentum_sum, sub.velocity_sum,
                             sub.num_states)
        # ---- merge into the trajectory (:580-660) ----
        if div_s or term_s:
            proposal = Proposal(proposal.state, proposal.energy, proposal.weight,
                                _logaddexp(proposal.sum_log_p_accept, sub_prop.sum_log_p_accept))
        else:
            proposal = _sample(float(u_merge[e]), proposal, sub_prop, biased=True)
        left, right = (traj, sub) if direction > 0 else (sub, traj)
        traj = Trajectory(left.leftmost_state, right.rightmost_state, left.momentum_sum + right.momentum_sum,
                          left.velocity_sum + right.velocity_sum, left.num_states + right.num_states)
        turn_full = _is_turning(criterion, traj.leftmost_state.momentum[0], traj.rightmost_state.momentum[0],
                                traj.momentum_sum, traj.leftmost_state.velocity[0],
                                traj.rightmost_state.velocity[0], traj.velocity_sum)
        e += 1
        diverging = div_s
        turning = term_s or turn_full
    return proposal, traj, e, diverging, turning


def build_kernel(divergence_threshold=1000.0, **solver_kwargs):
    """geomjax/rmhmc/nuts.py:78-161 with the keys replaced by
    noise = (z [C, D], u_dir [C, M], u_merge [C, M], u_step [C, 2^M - 1])."""

    def kernel(noise, state, logdensity_fn, step_size, metric_fn, max_num_doublings=10, stopping_criterion="euc"):
        z, u_dir, u_merge, u_step = noise
        momentum_generator, kinetic_energy_fn, imvp = gaussian_riemannian(metric_fn)
        one_step = implicit_midpoint(logdensity_fn, kinetic_energy_fn, imvp, **solver_kwargs)
        energy = hmc_energy(kinetic_energy_fn)
        position, logdensity, logdensity_grad = state
        momentum = vmap(momentum_generator)(z, position)
        velocity = vmap(imvp)(position, momentum)
        C = position.shape[0]
        M = int(max_num_doublings)
        outs, infos = [], []
        for c in range(C):
            s0 = IntegratorState(position[c:c + 1], momentum[c:c + 1], velocity[c:c + 1], logdensity[c:c + 1],
                                 logdensity_grad[c:c + 1])
            H0 = float(energy(s0)[0])
            proposal, traj, e, diverging, turning = _expand(s0, H0, one_step, energy, step_size, u_dir[c],
                                                            u_merge[c], u_step[c], M, stopping_criterion,
                                                            divergence_threshold)
            acc = math.exp(proposal.sum_log_p_accept) / traj.num_states
            outs.append(proposal.state)
            infos.append(NUTSInfo(momentum[c], diverging, turning, proposal.energy, traj.leftmost_state,
                                  traj.rightmost_state, e, traj.num_states, acc))
        new_state = RMHMCState(torch.cat([s.position for s in outs]), torch.cat([s.logdensity for s in outs]),
                               torch.cat([s.logdensity_grad for s in outs]))
        return new_state, infos

    return kernel


def build_lmcmonge_kernel(divergence_threshold=1000.0, max_num_doublings=10, reference_faithful=True):
    """geomjax/lmcmonge/nuts.py:80-207,300-404: the same expansion over the
    LMC-Monge lan_integrator (oracle.lmcmonge) with the energy -l + K - J, noise
    as for ``build_kernel``.  Returns (LMCState, list of NUTSInfo) with NUTSInfo's
    first field holding the initial velocity, as the reference's does."""
    from . import lmcmonge as olm

    M = int(max_num_doublings)

    def kernel(noise, state, logdensity_fn, step_size, inverse_mass_matrix, alpha2, stopping_criterion="euc"):
        z, u_dir, u_merge, u_step = noise
        position, logdensity, logdensity_grad, volume_adjustment = state
        (velocity_generator, kinetic_energy, set_weighted_gradient, normalizing_constant,
         metric_vector_product) = olm.gaussian_riemannian(alpha2, inverse_mass_matrix, reference_faithful)
        integrator = olm.lan_integrator(logdensity_fn, set_weighted_gradient, normalizing_constant,
                                        metric_vector_product, reference_faithful)
        energy = olm.lmcmonge_energy(kinetic_energy)
        determinant_metric = normalizing_constant(alpha2, logdensity_grad)
        sdm = torch.sqrt(determinant_metric)[:, None]
        gnorm = logdensity_grad / sdm
        hvp_fn = vmap(lambda x, v: olm.hvp(logdensity_fn, x, v))
        dl_ig = set_weighted_gradient(gnorm)
        Hdl_ig = hvp_fn(position, dl_ig) / sdm
        ig_Hdl_ig = set_weighted_gradient(Hdl_ig)
        velocity = velocity_generator(z, position, alpha2, dl_ig)
        momentum = metric_vector_product(velocity, alpha2, logdensity_grad, determinant_metric)
        hvn = hvp_fn(position, velocity) / sdm
        J = torch.as_tensor(volume_adjustment, dtype=position.dtype).expand(position.shape[0])
        a2 = torch.as_tensor(alpha2, dtype=position.dtype)
        C = position.shape[0]
        outs, infos = [], []
        for c in range(C):
            r = slice(c, c + 1)
            s0 = olm.RiemannianIntegratorState(a2, position[r], momentum[r], velocity[r], logdensity[r], gnorm[r],
                                               dl_ig[r], Hdl_ig[r], ig_Hdl_ig[r], hvn[r], determinant_metric[r],
                                               J[r])
            H0 = float(energy(s0)[0])
            proposal, traj, e, diverging, turning = _expand(s0, H0, integrator, energy, step_size, u_dir[c],
                                                            u_merge[c], u_step[c], M, stopping_criterion,
                                                            divergence_threshold)
            acc = math.exp(proposal.sum_log_p_accept) / traj.num_states
            outs.append(proposal.state)
            infos.append(NUTSInfo(velocity[c], diverging, turning, proposal.energy, traj.leftmost_state,
                                  traj.rightmost_state, e, traj.num_states, acc))
        new_state = olm.LMCState(
            torch.cat([s.position for s in outs]), torch.cat([s.logdensity for s in outs]),
            torch.cat([s.logdensity_grad_norm * torch.sqrt(s.determinant_metric)[:, None] for s in outs]),
            torch.cat([s.volume_adjustment.reshape(1) for s in outs]))
        return new_state, infos

    return kernel


def build_lmcmc_kernel(divergence_threshold=1000.0):
    """geomjax/lmcmc/nuts.py:78-161,253-351: the same expansion over the
    general-metric lan_integrator (oracle.lmcmc) with the energy
    -l - log det G / 2 + v^T G v / 2 - J, noise as for ``build_kernel``.  Returns
    (LMCState, list of NUTSInfo) with NUTSInfo's first field holding the initial
    velocity, as the reference's does."""
    from . import lmcmc as olc

    def kernel(noise, state, logdensity_fn, step_size, metric_fn, max_num_doublings=10, stopping_criterion="euc"):
        z, u_dir, u_merge, u_step = noise
        (velocity_generator, kinetic_energy, omega_tilde_fn, grad_logdetmetric,
         metric_vector_product) = olc.gaussian_riemannian(metric_fn)
        integrator = olc.lan_integrator(logdensity_fn, omega_tilde_fn, grad_logdetmetric, metric_vector_product)
        energy = olc.lmc_energy(kinetic_energy)
        position, logdensity, logdensity_grad, volume_adjustment = state
        velocity = vmap(velocity_generator)(z, position)
        momentum = vmap(metric_vector_product)(position, velocity)
        J = torch.as_tensor(volume_adjustment, dtype=position.dtype).expand(position.shape[0])
        M = int(max_num_doublings)
        outs, infos = [], []
        for c in range(position.shape[0]):
            r = slice(c, c + 1)
            s0 = olc.IntegratorState(position[r], momentum[r], velocity[r], logdensity[r], logdensity_grad[r], J[r])
            H0 = float(energy(s0)[0])
            proposal, traj, e, diverging, turning = _expand(s0, H0, integrator, energy, step_size, u_dir[c],
                                                            u_merge[c], u_step[c], M, stopping_criterion,
                                                            divergence_threshold)
            acc = math.exp(proposal.sum_log_p_accept) / traj.num_states
            outs.append(proposal.state)
            infos.append(NUTSInfo(velocity[c], diverging, turning, proposal.energy, traj.leftmost_state,
                                  traj.rightmost_state, e, traj.num_states, acc))
        new_state = olc.LMCState(torch.cat([s.position for s in outs]), torch.cat([s.logdensity for s in outs]),
                                 torch.cat([s.logdensity_grad for s in outs]),
                                 torch.cat([s.volume_adjustment.reshape(1) for s in outs]))
        return new_state, infos

    return kernel
