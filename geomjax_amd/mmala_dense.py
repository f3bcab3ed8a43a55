"""Manifold MALA with a general dense user metric (geomjax/lmcmc/mmala.py with any
``metric_fn``).

The user's functions are evaluated and differentiated by ``torch.func`` on the
GPU (the metric's Jacobian stands where the reference's ``jax.jacfwd`` of the
inverse metric does); the per-chain linear algebra runs in
``gjx_dense_mmala_drift`` (Cholesky of the metric, its inverse, Gamma_i =
sum_j d(G^-1)_ij/dq_j = -(G^-1 h)_i with h_a = sum_bj dG_ab/dq_j (G^-1)_bj, the
drift G^-1 grad l + Gamma, log det G and the draw chol(G^-1) z).  float64,
D <= 64.  Draws as in ``lmcmc_dense``.
"""
import math

import torch

from . import _lib
from .lmcmc_dense import _check, _metric, _noise, _value_and_grad
from .random import HostNoise, PRNGKey

__all__ = ["init", "transition", "sample"]


def _drift(G, dG=None, g=None, z=None):
    C, D = G.shape[0], G.shape[1]
    dev = G.device
    vec = lambda: torch.empty(C, D, dtype=torch.float64, device=dev)
    logdet = torch.empty(C, dtype=torch.float64, device=dev)
    gamma = vec() if dG is not None else None
    drift = vec() if g is not None else None
    ciz = vec() if z is not None else None
    io = _lib.DenseMetricIO(num_chains=C, dim=D, step_size=0.0, metric=_lib.ptr(G), metric_jacobian=_lib.ptr(dG),
                            logdensity_grad=_lib.ptr(g), z=_lib.ptr(z), logdet=logdet.data_ptr(),
                            gamma=_lib.ptr(gamma), drift=_lib.ptr(drift), chol_inv_z=_lib.ptr(ciz))
    _lib.check(_lib.load().gjx_dense_mmala_drift(io, _lib.stream_handle(dev)), "gjx_dense_mmala_drift")
    return logdet, gamma, drift, ciz


def init(position, logdensity_fn, metric_fn):
    """geomjax/lmcmc/mmala.py:53-60: log density, gradient, metric and Gamma."""
    from .lmcmc_dense import resolve_metric_fn
    from .mmala import DiffusionState

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)
    single = position.dim() == 1
    q = (position[None] if single else position).contiguous()
    _check(q)
    l, g = _value_and_grad(logdensity_fn, q)
    G, dG = _metric(metric_fn, q)
    _, gamma, _, _ = _drift(G, dG)
    if single:
        return DiffusionState(position, l[0], g[0], G[0], gamma[0])
    return DiffusionState(position, l, g, G, gamma)


def transition(rng_key, state, logdensity_fn, step_size, metric_fn, reference_faithful=True,
               divergence_threshold=1000.0):
    """geomjax/lmcmc/mmala.py:115-148 for every chain.  The current state's metric
    and Gamma are taken from ``state`` (as the reference's transition energy does)."""
    from .mmala import DiffusionState, MMALAInfo
    from .lmcmc_dense import resolve_metric_fn

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)

    x = state.position.contiguous()
    _check(x)
    C, D = x.shape
    dev = x.device
    eps = float(step_size)
    z, u = _noise(rng_key, C, D, dev)
    l0 = state.logdensity.reshape(C).to(torch.float64)
    g0 = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
    G0 = state.metric.reshape(C, D, D).to(torch.float64).contiguous()
    gamma0 = state.Gamma.reshape(C, D).to(torch.float64)
    # mu(x) = x + eps (G^-1 grad l + Gamma); the state's Gamma, the drift kernel's G^-1 grad l
    logdet0, _, gi_g, s = _drift(G0, None, g0, z)
    x1 = x + eps * (gi_g + gamma0) + math.sqrt(2.0 * eps) * s
    norm = -0.5 * D * (math.log(2.0 * eps) + math.log(2.0 * math.pi))
    logq_fwd = -0.5 * torch.sum(z * z, dim=1) + 0.5 * logdet0 + norm
    l1, g1 = _value_and_grad(logdensity_fn, x1)
    G1, dG1 = _metric(metric_fn, x1)
    logdet1, gamma1, drift1, _ = _drift(G1, dG1, g1)
    r = x - (x1 + eps * drift1)
    rGr = torch.sum(r * torch.bmm(G1, r[:, :, None])[:, :, 0], dim=1)
    logq_bwd = -rGr / (4.0 * eps) + 0.5 * logdet1 + norm
    new_energy = -l0 - logq_fwd
    prev_energy = -l1 - logq_bwd
    delta = prev_energy - new_energy if reference_faithful else new_energy - prev_energy
    delta = torch.where(torch.isnan(delta), torch.full_like(delta, -float("inf")), delta)
    p_accept = torch.clamp(torch.exp(delta), max=1.0)
    accept = u < p_accept
    a1, a2 = accept[:, None], accept[:, None, None]
    proposal = DiffusionState(x1, l1, g1, G1, gamma1)
    new_state = DiffusionState(torch.where(a1, x1, x), torch.where(accept, l1, l0), torch.where(a1, g1, g0),
                               torch.where(a2, G1, G0), torch.where(a1, gamma1, gamma0))
    info = MMALAInfo(p_accept, accept, proposal, -delta > divergence_threshold, delta)
    return new_state, info


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples, *, reference_faithful=True,
           divergence_threshold=1000, store_samples=True, samples_out=None):
    """``num_samples`` transitions (a Python-driven loop; capture it with
    ``GraphSampler`` to remove the host dispatch)."""
    from .mmala import DiffusionState
    from .sampling import SampleResult

    q = state.position
    C, D = q.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    dev = q.device
    samples = samples_out if samples_out is not None else (
        torch.zeros(T, C, D, dtype=torch.float64, device=dev) if store_samples else None)
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
    s = DiffusionState(*state)
    for t in range(T):
        key = PRNGKey(rng_key.seed, rng_key.counter + t) if isinstance(rng_key, PRNGKey) else \
            HostNoise(rng_key.z[t], rng_key.u[t])
        s, info = transition(key, s, logdensity_fn, step_size, metric_fn, reference_faithful, divergence_threshold)
        if samples is not None:
            samples[t].copy_(s.position)
        acc[t] = info.acceptance_rate.float()
        n_acc += info.is_accepted.int()
    fev = torch.full((C,), 2 * T, dtype=torch.int64, device=dev)
    return SampleResult(s, samples, acc, n_acc, fev)
