"""Generalized LMC: persistent velocity and a non-reversible slice
(geomjax/lmcmc/glmc.py:33-186, geomjax/mcmc/proposal.py:262-277).

Each transition refreshes the velocity partially, v = sqrt(1 - alpha) v_prev +
sqrt(alpha) L^-T z (L L^T = G(q)), translates the slice variable,
s = ((s + 1 + delta + noise) mod 2) - 1, takes one lan_integrator step and
accepts when log|s| <= H0 - H1; an accepted slice moves to s e^{-(H0 - H1)}.
After the step the velocity is flipped twice (lmc_proposal, then glmc): an
accepted move keeps the end velocity, a rejected one reverses the initial
velocity.  The returned ``LMCInfo.acceptance_rate`` carries the new slice, as
the reference's proposal does (proposal.py:274-275 feeding lmc.py:460-466).

The reference builds the six-field IntegratorState of
geomjax/lmcmc/integrators.py:26-38 from five values (glmc.py:170-172), a
TypeError at the first transition; here the state carries the momentum
G(q) v, as the lmcmc kernel builds it (geomjax/lmcmc/lmc.py:170-174).

Any metric: the user's ``logdensity_fn`` / ``metric_fn`` (torch functions)
are evaluated and differentiated by torch.func on the GPU and the metric
algebra (Cholesky velocity draw, log det G, the Christoffel contractions, LU
solves and volume terms) runs in the dense-metric HIP kernels, as for lmcmc
with a user metric (float64, D <= 64).  ``LogisticRegression`` with
``FisherMetric`` is taken as the pair of functions it describes.

Keys: ``HostNoise(z, u)`` -- z [C, D] the standard normals of the velocity
refresh (of the initial velocity in ``init``), u [C] uniforms (the slice in
``init``: -1 + 2u; in a transition the draw handed to ``noise_fn``) -- or
``PRNGKey`` (torch's device generator seeded from (seed, counter)).
"""
import math
from typing import NamedTuple

import torch

from .lmcmc_dense import (_check, _metric, _noise, _position, _value_and_grad, as_batch, lmc_step,
                          resolve_metric_fn, unbatch)
from .proposal import Proposal

__all__ = ["GLMCState", "init", "build_kernel", "glmc"]


class GLMCState(NamedTuple):  # geomjax/lmcmc/glmc.py:33-55 (same fields and order)
    position: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    slice: torch.Tensor
    volume_adjustment: torch.Tensor


def _functions(logdensity_fn, metric_fn):
    """The (log density, metric) pair as functions: a LogisticRegression with a
    FisherMetric becomes the model's log density and Fisher metric in torch."""
    from .metrics import FisherMetric
    from .models import LogisticRegression

    if isinstance(logdensity_fn, LogisticRegression):
        if metric_fn is not None and not isinstance(metric_fn, FisherMetric):
            raise NotImplementedError("glmc on a LogisticRegression descriptor takes metric_fn=FisherMetric()")
        m = logdensity_fn
        X, y, pv = m.X, m.y, m.prior_var

        def logp(b):
            eta = X @ b
            return (torch.sum(y * eta - torch.nn.functional.softplus(eta)) - 0.5 * torch.dot(b, b) / pv
                    - 0.5 * b.shape[0] * math.log(2.0 * math.pi * pv))

        def fisher(b):
            s = torch.sigmoid(X @ b)
            return (X.T * (s * (1.0 - s))) @ X + torch.eye(b.shape[0], dtype=b.dtype, device=b.device) / pv

        return logp, fisher
    if not callable(logdensity_fn):
        raise NotImplementedError(f"glmc takes a callable log density or a LogisticRegression "
                                  f"(got {logdensity_fn!r})")
    return logdensity_fn, (None if metric_fn is None else resolve_metric_fn(logdensity_fn, metric_fn))


def init(position, rng_key, logdensity_fn):
    """geomjax/lmcmc/glmc.py:58-71: log density and gradient, a standard normal
    velocity, a slice uniform on [-1, 1), no volume adjustment."""
    logp, _ = _functions(logdensity_fn, None)
    single = position.dim() == 1
    q = (position[None] if single else position).contiguous()
    _check(q)
    C, D = q.shape
    z, u = _noise(rng_key, C, D, q.device)
    l, g = _value_and_grad(logp, q)
    st = GLMCState(q, z, l, g, -1.0 + 2.0 * u, torch.zeros_like(l))
    return unbatch(st) if single else st


def transition(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta, noise_fn=None,
               divergence_threshold=1000.0):
    """One glmc transition of every chain (geomjax/lmcmc/glmc.py:123-169)."""
    from .lmcmc import IntegratorState, LMCInfo

    if not 0.0 <= float(alpha) <= 1.0:
        raise ValueError(f"alpha must be in [0, 1] (got {alpha})")
    logp, metric = _functions(logdensity_fn, metric_fn)
    q0 = state.position.contiguous()
    _check(q0)
    C, D = q0.shape
    dev = q0.device
    f64 = dict(dtype=torch.float64, device=dev)
    v_prev = state.velocity.reshape(C, D).to(**f64)
    l0 = state.logdensity.reshape(C).to(**f64)
    g0 = state.logdensity_grad.reshape(C, D).to(**f64).contiguous()
    s_prev = torch.as_tensor(state.slice, **f64).expand(C)
    J0 = torch.as_tensor(state.volume_adjustment, **f64).expand(C).clone()
    z, u = _noise(rng_key, C, D, dev)

    # partial refresh of the persistent velocity (glmc.py:172-190)
    G, dG = _metric(metric, q0)
    logdet, _, fresh, _ = _position(G, dG, z=z)
    v0 = math.sqrt(1.0 - float(alpha)) * v_prev + math.sqrt(float(alpha)) * fresh
    _, _, _, vgv0 = _position(G, dG, velocity=v0.contiguous())
    H0 = -l0 + (-0.5 * logdet + 0.5 * vgv0) - J0
    # the slice moves by a fixed amount plus the noise (glmc.py:156)
    noise = 0.0 if noise_fn is None else noise_fn(u)
    slice_ = torch.remainder(s_prev + 1.0 + float(delta) + noise, 2.0) - 1.0
    # one lan_integrator step (lmc_proposal with num_integration_steps = 1)
    eps = torch.full((C,), float(step_size), **f64)
    (q, momentum, v, l, g, J), H1 = lmc_step((q0, None, v0.contiguous(), l0, g0, J0), eps, logp, metric)
    dE = H0 - H1
    dE = torch.where(torch.isnan(dE), torch.full_like(dE, -float("inf")), dE)
    accept = torch.log(torch.abs(slice_)) <= dE
    new_slice = torch.where(accept, slice_ * torch.exp(-dE), slice_)
    a1 = accept[:, None]
    # flipped twice: the end velocity when accepted, the reversed initial one when rejected
    new_state = GLMCState(torch.where(a1, q, q0), torch.where(a1, v, -v0), torch.where(accept, l, l0),
                          torch.where(a1, g, g0), new_slice, torch.where(accept, J, J0))
    end = IntegratorState(q, -momentum, -v, l, g, J)  # the proposal carries the flipped end (lmc.py:478-499)
    proposal = Proposal(end, H1, dE, torch.minimum(dE, torch.zeros_like(dE)))
    info = LMCInfo(v0, new_slice, accept, -dE > divergence_threshold, H1, proposal, 1)
    return new_state, info


def build_kernel(noise_fn=None, divergence_threshold: float = 1000):
    """geomjax/lmcmc/glmc.py:74-169.  ``noise_fn`` maps the transition's [C]
    uniforms to the slice noise (default none: the reference's single atom at 0)."""

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta):
        state, single = as_batch(state)
        out = transition(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta, noise_fn,
                         divergence_threshold)
        return unbatch(out) if single else out

    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta, num_samples, *, noise_fn=None,
           divergence_threshold=1000, store_samples=True, samples_out=None):
    """``num_samples`` transitions of every chain (a loop of kernel launches);
    with a PRNGKey transition t uses counter + t."""
    from .random import HostNoise, PRNGKey
    from .sampling import SampleResult

    C, D = state.position.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    dev = state.position.device
    samples = samples_out if samples_out is not None else (
        torch.zeros(T, C, D, dtype=torch.float64, device=dev) if store_samples else None)
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
    s = state
    for t in range(T):
        key = PRNGKey(rng_key.seed, rng_key.counter + t) if isinstance(rng_key, PRNGKey) else \
            HostNoise(rng_key.z[t], rng_key.u[t])
        s, info = transition(key, s, logdensity_fn, step_size, metric_fn, alpha, delta, noise_fn,
                             divergence_threshold)
        if samples is not None:
            samples[t].copy_(s.position)
        acc[t].copy_(info.is_accepted.float())
        n_acc += info.is_accepted.to(torch.int32)
    fev = torch.full((C,), T, dtype=torch.int64, device=dev)
    return SampleResult(s, samples, acc, n_acc, fev)


class glmc:
    """geomjax/lmcmc/glmc.py:193-283: ``glmc(logdensity_fn, step_size, metric_fn,
    alpha, delta)`` with ``init(position, rng_key)`` and ``step(rng_key, state)``."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, alpha, delta, *, divergence_threshold: int = 1000,
                noise_gn=None):
        from .sampling import BoundAlgorithm

        kernel = cls.build_kernel(noise_gn, divergence_threshold)

        def init_fn(position, rng_key):
            return cls.init(position, rng_key, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta)

        def sample_fn(rng_key, state, num_samples, **kw):
            return sample(rng_key, state, logdensity_fn, step_size, metric_fn, alpha, delta, num_samples,
                          noise_fn=noise_gn, divergence_threshold=divergence_threshold, **kw)

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
