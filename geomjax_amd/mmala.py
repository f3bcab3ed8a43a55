"""Manifold MALA (geomjax/lmcmc/mmala.py) on the HIP kernel.

Same API as the reference: ``mmala(logdensity_fn, step_size, metric_fn)`` with
``init(position, logdensity_fn, metric_fn)`` returning ``DiffusionState(position,
logdensity, logdensity_grad, metric, Gamma)`` and ``step``; the kernel
``build_kernel()(rng_key, state, logdensity_fn, step_size, metric_fn)`` returns
``(DiffusionState, MMALAInfo)``.  One transition is the Euler step of the
Riemannian overdamped Langevin diffusion,
``x' = x + eps G^-1 grad l + eps Gamma + chol(2 eps G^-1) z``, with the
asymmetric Metropolis test; it runs in one HIP kernel, one workgroup per chain,
float64: Bayesian logistic regression with the Fisher metric, one posterior or a
``LogisticBatch``, every posterior with D <= 64.

``reference_faithful=True`` (the default) keeps the reference's pairing of the
transition energies, whose proposal weight is minus the log Metropolis-Hastings
ratio (geomjax/lmcmc/mmala.py:74-113 with geomjax/mcmc/proposal.py:151-158);
``False`` uses the ratio itself.  A plain callable ``metric_fn`` (with a
plain callable log density) runs the dense-metric path (``mmala_dense``).
``rng_key`` is a ``PRNGKey`` (device Philox)
or ``HostNoise(z [C, D], u [C])`` (the proposal's normals and the accept
uniform).
"""
from typing import NamedTuple

import torch

from . import _lib
from .metrics import FisherMetric
from .models import LogisticBatch, LogisticRegression, logistic_params
from .random import HostNoise, PRNGKey

__all__ = ["DiffusionState", "MMALAInfo", "init", "build_kernel", "sample", "mmala"]


class DiffusionState(NamedTuple):  # geomjax/lmcmc/diffusions.py:27-32 (same fields and order)
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    metric: torch.Tensor   # [C, D, D]
    Gamma: torch.Tensor    # [C, D]


class MMALAInfo(NamedTuple):  # geomjax/lmcmc/mmala.py:32-50 (same fields and order)
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    proposal: DiffusionState
    is_divergent: torch.Tensor
    weight: torch.Tensor


def _check(logdensity_fn, metric_fn, position):
    if not isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) or not isinstance(metric_fn, FisherMetric):
        raise NotImplementedError(f"no HIP MMALA kernel for logdensity_fn={logdensity_fn!r}, metric_fn={metric_fn!r}")
    if not position.is_cuda or position.dtype != torch.float64:
        raise ValueError("MMALA states must be float64 tensors on a HIP device")
    if position.shape[-1] != logdensity_fn.dim:
        raise ValueError(f"position has {position.shape[-1]} coordinates, the model has {logdensity_fn.dim}")


def init(position, logdensity_fn, metric_fn=FisherMetric()):
    """geomjax/lmcmc/mmala.py:53-60: log density, gradient, metric and Gamma.
    Plain callables (a user log density and metric) take the dense-metric path."""
    from .lmcmc_dense import is_dense_metric_fn

    if is_dense_metric_fn(logdensity_fn, metric_fn):
        from . import mmala_dense

        return mmala_dense.init(position, logdensity_fn, metric_fn)
    single = position.dim() == 1
    pos = (position[None] if single else position).contiguous()
    _check(logdensity_fn, metric_fn, pos)
    dev = pos.device
    C, D = pos.shape
    logp = torch.empty(C, dtype=torch.float64, device=dev)
    grad = torch.zeros(C, D, dtype=torch.float64, device=dev)
    metric = torch.zeros(C, D, D, dtype=torch.float64, device=dev)
    gamma = torch.zeros(C, D, dtype=torch.float64, device=dev)
    params = logistic_params(logdensity_fn, C)
    rc = _lib.load().gjx_mmala_fisher_logistic_init(params, logdensity_fn.data(), pos.data_ptr(), logp.data_ptr(),
                                                    grad.data_ptr(), metric.data_ptr(), gamma.data_ptr(),
                                                    _lib.stream_handle(dev))
    _lib.check(rc, "gjx_mmala_fisher_logistic_init")
    if single:
        return DiffusionState(position, logp[0], grad[0], metric[0], gamma[0])
    return DiffusionState(position, logp, grad, metric, gamma)


def _noise(rng_key, T, C, D, dev):
    if isinstance(rng_key, HostNoise):
        shape = (C, D) if T is None else (T, C, D)
        ushape = (C,) if T is None else (T, C)
        return (rng_key.z.reshape(shape).to(device=dev, dtype=torch.float64).contiguous(),
                rng_key.u.reshape(ushape).to(device=dev, dtype=torch.float64).contiguous(), 0, 0)
    if isinstance(rng_key, PRNGKey):
        return None, None, rng_key.seed, rng_key.counter
    raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")


def build_kernel(*, reference_faithful=True, divergence_threshold: float = 1000):
    """geomjax/lmcmc/mmala.py:63-148 (the reference fixes the divergence
    threshold at 1000)."""

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn):
        from .lmcmc_dense import is_dense_metric_fn

        if is_dense_metric_fn(logdensity_fn, metric_fn):
            from . import mmala_dense
            from .lmcmc_dense import as_batch, unbatch

            state, single = as_batch(state)
            out = mmala_dense.transition(rng_key, state, logdensity_fn, step_size, metric_fn, reference_faithful,
                                         divergence_threshold)
            return unbatch(out) if single else out
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        _check(logdensity_fn, metric_fn, position)
        dev = position.device
        C, D = position.shape
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        grad = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
        z, u, seed, counter = _noise(rng_key, None, C, D, dev)

        def vec():
            return torch.zeros(C, D, dtype=torch.float64, device=dev)

        def mat():
            return torch.zeros(C, D, D, dtype=torch.float64, device=dev)

        def sca(dt=torch.float64):
            return torch.empty(C, dtype=dt, device=dev)

        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), new_metric=mat(),
                   new_gamma=vec(), acceptance_rate=sca(), is_accepted=sca(torch.uint8),
                   is_divergent=sca(torch.uint8), weight=sca(), proposal_position=vec(), proposal_logdensity=sca(),
                   proposal_logdensity_grad=vec(), proposal_metric=mat(), proposal_gamma=vec())
        io = _lib.MmalaStepIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                              logdensity_grad=grad.data_ptr(), z=_lib.ptr(z), u=_lib.ptr(u),
                              **{k: v.data_ptr() for k, v in out.items()})
        params = logistic_params(logdensity_fn, C, step_size=step_size, divergence_threshold=divergence_threshold,
                                 seed=seed, counter=counter)
        rc = _lib.load().gjx_mmala_fisher_logistic_step(params, logdensity_fn.data(), int(bool(reference_faithful)),
                                                        io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_mmala_fisher_logistic_step")

        def sq(t):
            return t[0] if single else t

        new_state = DiffusionState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]),
                                   sq(out["new_metric"]), sq(out["new_gamma"]))
        proposal = DiffusionState(sq(out["proposal_position"]), sq(out["proposal_logdensity"]),
                                  sq(out["proposal_logdensity_grad"]), sq(out["proposal_metric"]),
                                  sq(out["proposal_gamma"]))
        info = MMALAInfo(sq(out["acceptance_rate"]), sq(out["is_accepted"].bool()), proposal,
                         sq(out["is_divergent"].bool()), sq(out["weight"]))
        return new_state, info

    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples, *, reference_faithful=True,
           divergence_threshold=1000, store_samples=True, samples_out=None):
    """``num_samples`` MMALA transitions of every chain in one persistent launch;
    the final state's gradient, metric and Gamma are evaluated at its position."""
    from .lmcmc_dense import is_dense_metric_fn
    from .sampling import SampleResult

    if is_dense_metric_fn(logdensity_fn, metric_fn):
        from . import mmala_dense

        return mmala_dense.sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples,
                                  reference_faithful=reference_faithful, divergence_threshold=divergence_threshold,
                                  store_samples=store_samples, samples_out=samples_out)
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
    z, u, seed, counter = _noise(rng_key, T, C, D, dev)
    final_q = torch.zeros_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.zeros_like(position)
    final_m = torch.zeros(C, D, D, dtype=torch.float64, device=dev)
    final_gamma = torch.zeros_like(position)
    if samples_out is not None:
        samples = samples_out
    else:
        samples = torch.zeros(T, C, D, dtype=torch.float64, device=dev) if store_samples else None
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.empty(C, dtype=torch.int32, device=dev)
    io = _lib.MmalaSampleIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(), z=_lib.ptr(z),
                            u=_lib.ptr(u), final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                            final_logdensity_grad=final_g.data_ptr(), final_metric=final_m.data_ptr(),
                            final_gamma=final_gamma.data_ptr(), samples=_lib.ptr(samples),
                            acceptance_rate=acc.data_ptr(), num_accepted=n_acc.data_ptr())
    params = logistic_params(logdensity_fn, C, T=T, step_size=step_size, divergence_threshold=divergence_threshold,
                             seed=seed, counter=counter)
    rc = _lib.load().gjx_mmala_fisher_logistic_sample(params, logdensity_fn.data(), int(bool(reference_faithful)),
                                                      io, _lib.stream_handle(dev))
    _lib.check(rc, "gjx_mmala_fisher_logistic_sample")
    fev = torch.full((C,), 2 * T, dtype=torch.int64, device=dev)  # two states evaluated per transition
    return SampleResult(DiffusionState(final_q, final_l, final_g, final_m, final_gamma), samples, acc, n_acc, fev)


class mmala:
    """geomjax/lmcmc/mmala.py:151-210 (exported as ``mmala``)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, *, reference_faithful=True):
        from .sampling import BoundAlgorithm

        kernel = cls.build_kernel(reference_faithful=reference_faithful)

        def init_fn(position):
            return cls.init(position, logdensity_fn, metric_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn)

        def sample_fn(rng_key, state, num_samples, **kw):
            return sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples,
                          reference_faithful=reference_faithful, **kw)

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
