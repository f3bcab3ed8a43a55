"""Lagrangian Monte Carlo with a general metric (lmcmc), on the HIP kernels.

Mirrors geomjax/lmcmc/lmc.py:30-346 and geomjax/lmcmc/integrators.py:
``LMCState``, ``LMCInfo``, ``IntegratorState``, ``init``, ``build_kernel``
returning ``kernel(rng_key, state, logdensity_fn, step_size, metric_fn,
num_integration_steps)``, and the ``lmc`` algorithm class (exported as
``geomjax_amd.lmc`` like the reference's alias, geomjax/__init__.py:17).

Served problems: Bayesian logistic regression (one posterior or a batch of
posteriors, D <= 64) with ``FisherMetric``: float64 states run the float64
kernel, float32 states the tuned one (split-bf16 MFMA Grams, float64
log-determinants, volume change and energies).
"""
from typing import NamedTuple

import torch

from . import _lib
from .metrics import FisherMetric
from .models import LogisticBatch, LogisticRegression, logistic_params
from .proposal import Proposal
from .random import HostNoise, PRNGKey

__all__ = ["LMCState", "LMCInfo", "IntegratorState", "init", "build_kernel", "lmc", "lan_integrator", "sample",
           "DynamicLMCState", "init_dynamic", "build_dynamic_kernel", "dynamic_lmc"]


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
    if not position.is_cuda or position.dtype not in (torch.float64, torch.float32):
        raise ValueError("LMC states must be float64 (plain) or float32 (tuned) tensors on a HIP device")
    if position.shape[-1] != logdensity_fn.dim:
        raise ValueError(f"position has {position.shape[-1]} coordinates, the model has {logdensity_fn.dim}")
    if isinstance(logdensity_fn, LogisticBatch) and position.shape[0] != logdensity_fn.num_chains:
        raise ValueError(f"the batch has {logdensity_fn.num_chains} chains, got {position.shape[0]}")


def init(position, logdensity_fn):
    """geomjax/lmcmc/lmc.py:98-101.  A plain callable log density (a user
    function) is evaluated by torch.func on the device (the dense-metric path)."""
    if not isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)):
        if callable(logdensity_fn):
            from . import lmcmc_dense

            return lmcmc_dense.init(position, logdensity_fn)
        raise NotImplementedError(f"no HIP log density for {logdensity_fn!r}")
    logdensity, grad = logdensity_fn.value_and_grad(position)
    return LMCState(position, logdensity, grad, torch.zeros_like(logdensity))


def _noise(rng_key, T, C, D, dev, dtype=torch.float64):
    if isinstance(rng_key, HostNoise):
        shape = (C, D) if T is None else (T, C, D)
        ushape = (C,) if T is None else (T, C)
        return (rng_key.z.reshape(shape).to(device=dev, dtype=dtype).contiguous(),
                rng_key.u.reshape(ushape).to(device=dev, dtype=torch.float64).contiguous(), 0, 0)
    if isinstance(rng_key, PRNGKey):
        return None, None, rng_key.seed, rng_key.counter
    raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")


def build_kernel(integrator=lan_integrator, divergence_threshold: float = 1000):
    """geomjax/lmcmc/lmc.py:114-182."""
    if integrator is not lan_integrator:
        raise NotImplementedError("only lan_integrator is implemented on the GPU")

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, *,
               max_num_integration_steps=None):
        """``num_integration_steps``: an int, or a [C] integer tensor of per-chain
        counts (the dynamic kernel's form), bounded by ``max_num_integration_steps``."""
        from . import lmcmc_dense
        from .rmhmc import _steps

        if lmcmc_dense.is_dense_metric_fn(logdensity_fn, metric_fn):
            # a user metric_fn: torch.func derivatives + the dense-metric HIP kernels
            state, single = lmcmc_dense.as_batch(state)
            out = lmcmc_dense.transition(rng_key, state, logdensity_fn, step_size, metric_fn,
                                         num_integration_steps, divergence_threshold, max_num_integration_steps)
            return lmcmc_dense.unbatch(out) if single else out
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        _check(logdensity_fn, metric_fn, position)
        dev = position.device
        C, D = position.shape
        dt = position.dtype
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        grad = state.logdensity_grad.reshape(C, D).to(dt).contiguous()
        J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
        z, u, seed, counter = _noise(rng_key, None, C, D, dev, dt)
        L, steps = _steps(num_integration_steps, (C,), dev, max_num_integration_steps)

        def vec():
            return torch.zeros(C, D, dtype=dt, device=dev)

        def sca(dt=torch.float64):
            return torch.empty(C, dtype=dt, device=dev)

        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), new_volume_adjustment=sca(),
                   velocity=vec(), acceptance_rate=sca(), is_accepted=sca(torch.uint8),
                   is_divergent=sca(torch.uint8), energy=sca(), proposal_position=vec(), proposal_momentum=vec(),
                   proposal_velocity=vec(), proposal_logdensity=sca(), proposal_logdensity_grad=vec(),
                   proposal_volume_adjustment=sca(), proposal_weight=sca(), proposal_sum_log_p_accept=sca())
        io = _lib.LmcmcStepIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                              logdensity_grad=grad.data_ptr(), volume_adjustment=J.data_ptr(), z=_lib.ptr(z),
                              u=_lib.ptr(u), num_integration_steps=_lib.ptr(steps),
                              **{k: v.data_ptr() for k, v in out.items()})
        params = logistic_params(logdensity_fn, C, L=L, step_size=step_size,
                                 divergence_threshold=divergence_threshold, seed=seed, counter=counter, dtype=dt)
        data = logdensity_fn.data(dtype=dt, num_chains=C)
        rc = _lib.load().gjx_lmc_fisher_logistic_step(params, data, io, _lib.stream_handle(dev))
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
                       sq(out["is_divergent"].bool()), sq(out["energy"]), proposal,
                       L if steps is None else sq(steps))
        return new_state, info

    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples, *,
           divergence_threshold=1000, store_samples=True, samples_out=None, max_num_integration_steps=None):
    """``num_samples`` LMC transitions of every chain in one persistent launch
    (``num_integration_steps`` an int or a [T, C] integer tensor)."""
    from . import lmcmc_dense
    from .rmhmc import _steps
    from .sampling import SampleResult

    if lmcmc_dense.is_dense_metric_fn(logdensity_fn, metric_fn):
        return lmcmc_dense.sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps,
                                  num_samples, divergence_threshold=divergence_threshold,
                                  store_samples=store_samples, samples_out=samples_out,
                                  max_num_integration_steps=max_num_integration_steps)
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
    dt = position.dtype
    logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
    J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
    z, u, seed, counter = _noise(rng_key, T, C, D, dev, dt)
    L, steps = _steps(num_integration_steps, (T, C), dev, max_num_integration_steps)
    final_q = torch.zeros_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.zeros_like(position)
    final_J = torch.empty(C, dtype=torch.float64, device=dev)
    if samples_out is not None:
        samples = samples_out
    else:
        samples = torch.zeros(T, C, D, dtype=dt, device=dev) if store_samples else None
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.empty(C, dtype=torch.int32, device=dev)
    io = _lib.LmcmcSampleIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                            volume_adjustment=J.data_ptr(), z=_lib.ptr(z), u=_lib.ptr(u),
                            final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                            final_logdensity_grad=final_g.data_ptr(), final_volume_adjustment=final_J.data_ptr(),
                            samples=_lib.ptr(samples), acceptance_rate=acc.data_ptr(), num_accepted=n_acc.data_ptr(),
                            num_integration_steps=_lib.ptr(steps))
    params = logistic_params(logdensity_fn, C, L=L, T=T, step_size=step_size,
                             divergence_threshold=divergence_threshold, seed=seed, counter=counter, dtype=dt)
    data = logdensity_fn.data(dtype=dt, num_chains=C)
    rc = _lib.load().gjx_lmc_fisher_logistic_sample(params, data, io, _lib.stream_handle(dev))
    _lib.check(rc, "gjx_lmc_fisher_logistic_sample")
    if steps is None:
        fev = torch.full((C,), T * L, dtype=torch.int64, device=dev)
    else:
        fev = steps.clamp(0, L).to(torch.int64).sum(0)
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


# ---------------------------------------------------------------------------
# Dynamic LMC (geomjax/lmcmc/lmc.py:45-57,104-111,185-252,349-408)
# ---------------------------------------------------------------------------


class DynamicLMCState(NamedTuple):  # geomjax/lmcmc/lmc.py:45-57 (same fields and order)
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor
    random_generator_arg: torch.Tensor  # [C] int64, one per chain


def init_dynamic(position, logdensity_fn, random_generator_arg):
    """geomjax/lmcmc/lmc.py:104-111; ``random_generator_arg`` is a scalar or a [C]
    int64 tensor."""
    s = init(position, logdensity_fn)
    arg = torch.as_tensor(random_generator_arg, dtype=torch.int64, device=s.position.device)
    C = s.position.shape[0] if s.position.dim() == 2 else 1
    return DynamicLMCState(s.position, s.logdensity, s.logdensity_grad, s.volume_adjustment, arg.expand(C).clone())


def build_dynamic_kernel(integrator=lan_integrator, divergence_threshold: float = 1000, next_random_arg_fn=None,
                         integration_steps_fn=None, max_num_integration_steps=9):
    """geomjax/lmcmc/lmc.py:185-252: the LMC kernel with a number of integration
    steps drawn per chain from its ``random_generator_arg``."""
    from .rmhmc import _integration_steps, _next_random_arg

    next_random_arg_fn = next_random_arg_fn or _next_random_arg
    integration_steps_fn = integration_steps_fn or _integration_steps
    base = build_kernel(integrator, divergence_threshold)

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, **integration_steps_kwargs):
        steps = integration_steps_fn(state.random_generator_arg, **integration_steps_kwargs)
        lmc_state = LMCState(state.position, state.logdensity, state.logdensity_grad, state.volume_adjustment)
        new, info = base(rng_key, lmc_state, logdensity_fn, step_size, metric_fn, steps,
                         max_num_integration_steps=max_num_integration_steps)
        return (DynamicLMCState(new.position, new.logdensity, new.logdensity_grad, new.volume_adjustment,
                                next_random_arg_fn(state.random_generator_arg)), info)

    return kernel


class dynamic_lmc:
    """geomjax/lmcmc/lmc.py:349-408."""

    init = staticmethod(init_dynamic)
    build_kernel = staticmethod(build_dynamic_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, *, divergence_threshold: int = 1000,
                integrator=lan_integrator, next_random_arg_fn=None, integration_steps_fn=None,
                max_num_integration_steps=9):
        from .rmhmc import _integration_steps, _next_random_arg
        from .sampling import BoundAlgorithm

        next_random_arg_fn = next_random_arg_fn or _next_random_arg
        integration_steps_fn = integration_steps_fn or _integration_steps
        kernel = cls.build_kernel(integrator, divergence_threshold, next_random_arg_fn, integration_steps_fn,
                                  max_num_integration_steps)

        def init_fn(position, random_generator_arg):
            return cls.init(position, logdensity_fn, random_generator_arg)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn)

        def sample_fn(rng_key, state, num_samples, **kw):
            """``num_samples`` transitions in one persistent launch; the step counts
            of all transitions are drawn up front from the successive arguments."""
            args, a = [], state.random_generator_arg
            for _ in range(int(num_samples)):
                args.append(integration_steps_fn(a))
                a = next_random_arg_fn(a)
            r = sample(rng_key, LMCState(state.position, state.logdensity, state.logdensity_grad,
                                         state.volume_adjustment), logdensity_fn, step_size, metric_fn,
                       torch.stack(args), num_samples, divergence_threshold=divergence_threshold,
                       max_num_integration_steps=max_num_integration_steps, **kw)
            return r._replace(state=DynamicLMCState(r.state.position, r.state.logdensity, r.state.logdensity_grad,
                                                    r.state.volume_adjustment, a))

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
