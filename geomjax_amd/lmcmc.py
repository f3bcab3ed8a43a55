"""Lagrangian Monte Carlo with a general metric (lmcmc), on the HIP kernels.

Mirrors geomjax/lmcmc/lmc.py:30-346 and geomjax/lmcmc/integrators.py:
``LMCState``, ``LMCInfo``, ``IntegratorState``, ``init``, ``build_kernel``
returning ``kernel(rng_key, state, logdensity_fn, step_size, metric_fn,
num_integration_steps)``, and the ``lmc`` algorithm class (exported as
``geomjax_amd.lmc`` like the reference's alias, geomjax/__init__.py:17).

Served problems: Bayesian logistic regression (one posterior or a batch of
posteriors) with ``FisherMetric``; float64.
"""
from typing import NamedTuple

import torch

from . import _lib
from .metrics import FisherMetric
from .models import LogisticBatch, LogisticRegression, logistic_params
from .proposal import Proposal
from .random import HostNoise, PRNGKey

__all__ = ["LMCState", "LMCInfo", "IntegratorState", "init", "build_kernel", "lmc", "lan_integrator", "sample"]


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


def lan_integrator(*args, **kwargs):
    """Marker for the explicit Lagrangian integrator (integrators.py:51-144); runs inside the HIP kernel."""
    raise TypeError("lan_integrator is executed inside the HIP kernel; pass it to build_kernel/lmc")


def _check(logdensity_fn, metric_fn, position):
    if not isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) or not isinstance(metric_fn, FisherMetric):
        raise NotImplementedError(f"no HIP LMC kernel for logdensity_fn={logdensity_fn!r}, metric_fn={metric_fn!r}")
    if not position.is_cuda or position.dtype != torch.float64:
        raise ValueError("LMC states must be float64 tensors on a HIP device")
    if position.shape[-1] != logdensity_fn.dim:
        raise ValueError(f"position has {position.shape[-1]} coordinates, the model has {logdensity_fn.dim}")
    if isinstance(logdensity_fn, LogisticBatch) and position.shape[0] != logdensity_fn.num_chains:
        raise ValueError(f"the batch has {logdensity_fn.num_chains} chains, got {position.shape[0]}")


def init(position, logdensity_fn):
    """geomjax/lmcmc/lmc.py:98-101."""
    if not isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)):
        raise NotImplementedError(f"no HIP log density for {logdensity_fn!r}")
    logdensity, grad = logdensity_fn.value_and_grad(position)
    return LMCState(position, logdensity, grad, torch.zeros_like(logdensity))


def _noise(rng_key, T, C, D, dev):
    if isinstance(rng_key, HostNoise):
        shape = (C, D) if T is None else (T, C, D)
        ushape = (C,) if T is None else (T, C)
        return (rng_key.z.reshape(shape).to(device=dev, dtype=torch.float64).contiguous(),
                rng_key.u.reshape(ushape).to(device=dev, dtype=torch.float64).contiguous(), 0, 0)
    if isinstance(rng_key, PRNGKey):
        return None, None, rng_key.seed, rng_key.counter
    raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")


def build_kernel(integrator=lan_integrator, divergence_threshold: float = 1000):
    """geomjax/lmcmc/lmc.py:114-182."""
    if integrator is not lan_integrator:
        raise NotImplementedError("only lan_integrator is implemented on the GPU")

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps):
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        _check(logdensity_fn, metric_fn, position)
        dev = position.device
        C, D = position.shape
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        grad = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
        J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
        z, u, seed, counter = _noise(rng_key, None, C, D, dev)
        L = int(num_integration_steps)

        def vec():
            return torch.zeros(C, D, dtype=torch.float64, device=dev)

        def sca(dt=torch.float64):
            return torch.empty(C, dtype=dt, device=dev)

        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), new_volume_adjustment=sca(),
                   velocity=vec(), acceptance_rate=sca(), is_accepted=sca(torch.uint8),
                   is_divergent=sca(torch.uint8), energy=sca(), proposal_position=vec(), proposal_momentum=vec(),
                   proposal_velocity=vec(), proposal_logdensity=sca(), proposal_logdensity_grad=vec(),
                   proposal_volume_adjustment=sca(), proposal_weight=sca(), proposal_sum_log_p_accept=sca())
        io = _lib.LmcmcStepIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                              logdensity_grad=grad.data_ptr(), volume_adjustment=J.data_ptr(), z=_lib.ptr(z),
                              u=_lib.ptr(u), **{k: v.data_ptr() for k, v in out.items()})
        params = logistic_params(logdensity_fn, C, L=L, step_size=step_size,
                                 divergence_threshold=divergence_threshold, seed=seed, counter=counter)
        rc = _lib.load().gjx_lmc_fisher_logistic_step(params, logdensity_fn.data(), io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_lmc_fisher_logistic_step")

        def sq(t):
            return t[0] if single else t

        new_state = LMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]),
                             sq(out["new_volume_adjustment"]))
        pstate = IntegratorState(sq(out["proposal_position"]), sq(out["proposal_momentum"]),
                                 sq(out["proposal_velocity"]), sq(out["proposal_logdensity"]),
                                 sq(out["proposal_logdensity_grad"]), sq(out["proposal_volume_adjustment"]))
        proposal = Proposal(pstate, sq(out["energy"]), sq(out["proposal_weight"]),
                            sq(out["proposal_sum_log_p_accept"]))
        info = LMCInfo(sq(out["velocity"]), sq(out["acceptance_rate"]), sq(out["is_accepted"].bool()),
                       sq(out["is_divergent"].bool()), sq(out["energy"]), proposal, L)
        return new_state, info

    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples, *,
           divergence_threshold=1000, store_samples=True, samples_out=None):
    """``num_samples`` LMC transitions of every chain in one persistent launch."""
    from .sampling import SampleResult

    position = state.position
    if position.dim() != 2:
        raise ValueError("sample() runs a batch of chains: position must be [C, D]")
    position = position.contiguous()
    _check(logdensity_fn, metric_fn, position)
    dev = position.device
    C, D = position.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
    J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
    z, u, seed, counter = _noise(rng_key, T, C, D, dev)
    final_q = torch.zeros_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.zeros_like(position)
    final_J = torch.empty(C, dtype=torch.float64, device=dev)
    if samples_out is not None:
        samples = samples_out
    else:
        samples = torch.zeros(T, C, D, dtype=torch.float64, device=dev) if store_samples else None
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.empty(C, dtype=torch.int32, device=dev)
    io = _lib.LmcmcSampleIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                            volume_adjustment=J.data_ptr(), z=_lib.ptr(z), u=_lib.ptr(u),
                            final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                            final_logdensity_grad=final_g.data_ptr(), final_volume_adjustment=final_J.data_ptr(),
                            samples=_lib.ptr(samples), acceptance_rate=acc.data_ptr(), num_accepted=n_acc.data_ptr())
    params = logistic_params(logdensity_fn, C, L=num_integration_steps, T=T, step_size=step_size,
                             divergence_threshold=divergence_threshold, seed=seed, counter=counter)
    rc = _lib.load().gjx_lmc_fisher_logistic_sample(params, logdensity_fn.data(), io, _lib.stream_handle(dev))
    _lib.check(rc, "gjx_lmc_fisher_logistic_sample")
    fev = torch.full((C,), T * int(num_integration_steps), dtype=torch.int64, device=dev)
    return SampleResult(LMCState(final_q, final_l, final_g, final_J), samples, acc, n_acc, fev)


class lmc:
    """geomjax/lmcmc/lmc.py:255-346 (exported as ``lmc``)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, num_integration_steps, *, divergence_threshold: int = 1000,
                integrator=lan_integrator):
        from .sampling import BoundAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps)

        def sample_fn(rng_key, state, num_samples, **kw):
            return sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples,
                          divergence_threshold=divergence_threshold, **kw)

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
