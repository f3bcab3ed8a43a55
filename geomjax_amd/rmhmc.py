"""RMHMC with the implicit midpoint integrator, on the HIP kernels.

Mirrors the reference's public surface (geomjax/rmhmc/rmhmc.py:30-311 and
geomjax/rmhmc/integrators.py:25-156): ``RMHMCState``, ``RMHMCInfo``,
``IntegratorState``, ``FixedPointIterationInfo``, ``init``, ``build_kernel``
returning ``kernel(rng_key, state, logdensity_fn, step_size, metric_fn,
num_integration_steps)``, and the ``rmhmc`` algorithm class whose
``__new__`` returns ``SamplingAlgorithm(init, step)``.  Keyword names are the
reference's, so callers that pass them by keyword (the adaptation drivers,
geomjax/adaptation/step_size_adaptation.py:147-153) work unchanged.

Differences forced by the GPU (documented, not silent):
  * states are batched: ``position`` is [C, D] (a [D] vector is one chain);
  * ``logdensity_fn`` / ``metric_fn`` are model and metric descriptors
    (``geomjax_amd.models``, ``geomjax_amd.metrics``) for the closed-form
    kernels; plain callables (a user log density and dense metric written with
    torch operations) run the dense-metric path (``rmhmc_dense``);
  * ``rng_key`` is a ``geomjax_amd.random.HostNoise`` or ``PRNGKey``.
The fixed-point solver's keywords (``max_iters``, ``convergence_tol``,
``divergence_tol``) are bound with ``functools.partial(implicit_midpoint,
...)`` exactly as with the reference.
"""
import functools
from typing import NamedTuple

import torch

from . import _lib
from .base import SamplingAlgorithm
from .metrics import FisherMetric, MongeMetric
from .models import LogisticBatch, LogisticRegression, NealFunnel, _dtype_code, logistic_params
from .proposal import Proposal
from .random import HostNoise, PRNGKey

__all__ = ["RMHMCState", "RMHMCInfo", "IntegratorState", "FixedPointIterationInfo", "init", "build_kernel",
           "rmhmc", "implicit_midpoint"]


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


_SOLVER_DEFAULTS = dict(convergence_tol=1e-6, divergence_tol=1e10, max_iters=100)


def implicit_midpoint(logdensity_fn=None, kinetic_energy_fn=None, inverse_metric_vector_product=None, *,
                      solver=None, return_info=False, **solver_kwargs):
    """Marker for the implicit midpoint integrator (integrators.py:92-156).

    The integration itself runs inside the HIP kernel; this function only
    exists so that ``functools.partial(implicit_midpoint, max_iters=8)`` binds
    solver keywords the way it does in the reference.
    """
    raise TypeError("implicit_midpoint is executed inside the HIP kernel; pass it to build_kernel/rmhmc")


def _solver_kwargs(integrator):
    kw = dict(_SOLVER_DEFAULTS)
    if integrator is implicit_midpoint:
        return kw
    if isinstance(integrator, functools.partial) and integrator.func is implicit_midpoint:
        extra = dict(integrator.keywords)
        for k in ("solver", "return_info"):
            extra.pop(k, None)
        unknown = set(extra) - set(_SOLVER_DEFAULTS)
        if unknown:
            raise TypeError(f"unknown solver keywords {sorted(unknown)}")
        kw.update(extra)
        return kw
    raise NotImplementedError("only implicit_midpoint (optionally functools.partial-bound) is implemented on the GPU")


def _as_batch(t):
    return (t[None], True) if t.dim() == 1 else (t, False)


def _steps(num_integration_steps, shape, dev, max_num_integration_steps=None):
    """(L, per-chain tensor or None): an int gives every chain L steps; an integer
    tensor of ``shape`` gives each chain (and transition) its own count, the
    dynamic kernel's form (geomjax/rmhmc/rmhmc.py:179-244), with L its maximum
    (``max_num_integration_steps`` if given, which avoids reading the tensor
    back to the host)."""
    if isinstance(num_integration_steps, torch.Tensor):
        t = num_integration_steps.to(device=dev, dtype=torch.int32).reshape(shape).contiguous()
        L = int(max_num_integration_steps) if max_num_integration_steps is not None else int(t.max().item())
        if L < 0:
            raise ValueError("num_integration_steps must be >= 0")
        return max(L, 0), t
    return int(num_integration_steps), None


def init(position, logdensity_fn):
    """geomjax/rmhmc/rmhmc.py:96-98: value and gradient of the log density.  A
    plain callable (a user function) is evaluated by torch.func on the device."""
    if not isinstance(logdensity_fn, (NealFunnel, LogisticRegression, LogisticBatch)):
        if callable(logdensity_fn):
            from . import rmhmc_dense

            return rmhmc_dense.init(position, logdensity_fn)
        raise NotImplementedError(f"no HIP log density for {logdensity_fn!r}")
    logdensity, grad = logdensity_fn.value_and_grad(position)
    return RMHMCState(position, logdensity, grad)


def _problem(logdensity_fn, metric_fn, position):
    """Which native kernel serves (logdensity_fn, metric_fn); validates shapes."""
    if not position.is_cuda:
        raise ValueError("RMHMC states must live on a HIP device (cuda:N)")
    if isinstance(logdensity_fn, NealFunnel) and isinstance(metric_fn, MongeMetric):
        if position.shape[-1] != logdensity_fn.D:
            raise ValueError(f"position has {position.shape[-1]} coordinates, the model has D={logdensity_fn.D}")
        return "funnel_monge"
    if isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) and isinstance(metric_fn, FisherMetric):
        if position.shape[-1] != logdensity_fn.dim:
            raise ValueError(f"position has {position.shape[-1]} coordinates, the model has {logdensity_fn.dim}")
        if position.dtype not in (torch.float64, torch.float32):
            raise TypeError("logistic RMHMC states are float64 (plain) or float32 (tuned)")
        if isinstance(logdensity_fn, LogisticBatch) and position.shape[0] != logdensity_fn.num_chains:
            raise ValueError(f"the batch has {logdensity_fn.num_chains} chains, got {position.shape[0]}")
        return "logistic_fisher"
    raise NotImplementedError(f"no HIP RMHMC kernel for logdensity_fn={logdensity_fn!r}, metric_fn={metric_fn!r}")


def _params(C, D, dtype_code, step_size, L, solver, alpha2, sigma, divergence_threshold, T=1,
            lanes_per_chain=0, seed=0, counter=0):
    return _lib.FunnelParams(
        num_chains=C, dim=D, dtype=dtype_code, num_integration_steps=int(L),
        max_iters=int(solver["max_iters"]), num_transitions=int(T), lanes_per_chain=int(lanes_per_chain),
        flags=0, step_size=float(step_size), alpha2=float(alpha2), sigma=float(sigma),
        convergence_tol=float(solver["convergence_tol"]), divergence_tol=float(solver["divergence_tol"]),
        divergence_threshold=float(divergence_threshold), seed=int(seed), counter=int(counter))


def _lanes(lanes_per_chain, kind, D, C, dtype):
    """Lanes per chain of a funnel kernel: the caller's, else the launch shape
    the autotuner measured for (kernel, D, chains, dtype), else 0 (the
    library's default)."""
    if lanes_per_chain:
        return int(lanes_per_chain)
    from .autotune import lookup

    return lookup(kind, D, C, dtype) or 0


def _waves(waves_per_chain, model, C, dtype, step_size=0.05, L=4, max_iters=6):
    """The C ABI's waves_per_chain field for a logistic call: an explicit 1 or
    2 applies to both one-workgroup widths (up to 32 and up to 64 features),
    a pair (w32, w64) to each; 0 takes the autotuner's choice for this problem
    and chain count (tuned on first use with the call's step size, number of
    steps and fixed-point cap), or the library default.  float64 calls ignore
    it."""
    if dtype != torch.float32:
        return 0
    if isinstance(waves_per_chain, tuple):
        a, b = (int(x) for x in waves_per_chain)
        if a not in (0, 1, 2) or b not in (0, 1, 2):
            raise ValueError(f"waves_per_chain entries must be 0, 1 or 2 (got {waves_per_chain})")
        return a | (b << 4)
    if waves_per_chain:
        if waves_per_chain not in (1, 2):
            raise ValueError(f"waves_per_chain must be 0, 1 or 2 (got {waves_per_chain})")
        return int(waves_per_chain) | (int(waves_per_chain) << 4)
    from .autotune import logistic_waves_field

    return logistic_waves_field(model, C, step_size=float(step_size), num_integration_steps=int(L),
                                max_iters=int(max_iters))


def build_kernel(integrator=implicit_midpoint, divergence_threshold: float = 1000, *,
                 return_fixed_point_info=False, lanes_per_chain=0, waves_per_chain=0):
    """geomjax/rmhmc/rmhmc.py:110-176.

    With ``return_fixed_point_info=True`` the kernel returns a third element,
    ``FixedPointIterationInfo`` of shape [C, num_integration_steps] (the info
    the reference computes at integrators.py:88 and discards at :153).
    """
    solver = _solver_kwargs(integrator)

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, *,
               max_num_integration_steps=None):
        from .lmcmc_dense import is_dense_metric_fn

        if is_dense_metric_fn(logdensity_fn, metric_fn):
            # a user metric_fn: torch.func derivatives + the dense-metric field kernel
            from . import rmhmc_dense
            from .lmcmc_dense import as_batch, unbatch

            state, single = as_batch(state)
            new, info, fps = rmhmc_dense.transition(rng_key, state, logdensity_fn, step_size, metric_fn,
                                                    num_integration_steps, solver, divergence_threshold,
                                                    max_num_integration_steps)
            out = (new, info)
            if return_fixed_point_info:
                C = state.position.shape[0]
                if fps:
                    fp = FixedPointIterationInfo(torch.stack([f[0] for f in fps], 1), torch.stack([f[1] for f in fps], 1),
                                                 torch.stack([f[2] for f in fps], 1).to(torch.int32))
                else:
                    dev = state.position.device
                    fp = FixedPointIterationInfo(torch.zeros(C, 0, dtype=torch.bool, device=dev),
                                                 torch.zeros(C, 0, dtype=torch.float64, device=dev),
                                                 torch.zeros(C, 0, dtype=torch.int32, device=dev))
                out = (new, info, fp)
            return unbatch(out) if single else out
        position, single = _as_batch(state.position)
        kind = _problem(logdensity_fn, metric_fn, position)
        position = position.contiguous()
        dev = position.device
        C, D = position.shape
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        grad = state.logdensity_grad.reshape(C, D).to(position.dtype).contiguous()
        L, steps = _steps(num_integration_steps, (C,), dev, max_num_integration_steps)
        if isinstance(rng_key, HostNoise):
            z = rng_key.z.reshape(C, D).to(device=dev, dtype=position.dtype).contiguous()
            u = rng_key.u.reshape(C).to(device=dev, dtype=torch.float64).contiguous()
            seed, counter = 0, 0
        elif isinstance(rng_key, PRNGKey):
            z = u = None
            seed, counter = rng_key.seed, rng_key.counter
        else:
            raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")

        def vec():
            return torch.empty(C, D, dtype=position.dtype, device=dev)

        def sca(dt=torch.float64):
            return torch.empty(C, dtype=dt, device=dev)

        out = dict(
            new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), momentum=vec(),
            acceptance_rate=sca(), is_accepted=sca(torch.uint8), is_divergent=sca(torch.uint8), energy=sca(),
            proposal_position=vec(), proposal_momentum=vec(), proposal_velocity=vec(), proposal_logdensity=sca(),
            proposal_logdensity_grad=vec(), proposal_weight=sca(), proposal_sum_log_p_accept=sca(),
            fp_iters=torch.empty(C, max(L, 1), dtype=torch.int32, device=dev),
            fp_norm=torch.empty(C, max(L, 1), dtype=torch.float64, device=dev),
            fp_success=torch.empty(C, max(L, 1), dtype=torch.uint8, device=dev),
        )
        io = _lib.RmhmcStepIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                              logdensity_grad=grad.data_ptr(), z=_lib.ptr(z), u=_lib.ptr(u),
                              num_integration_steps=_lib.ptr(steps),
                              **{k: v.data_ptr() for k, v in out.items()})
        if kind == "funnel_monge":
            params = _params(C, D, _dtype_code(position), step_size, L, solver, metric_fn.alpha2,
                             logdensity_fn.sigma, divergence_threshold, seed=seed,
                             lanes_per_chain=_lanes(lanes_per_chain, "rmhmc_monge_funnel", D, C, position.dtype),
                             counter=counter)
            rc = _lib.load().gjx_rmhmc_monge_funnel_step(params, io, _lib.stream_handle(dev))
            _lib.check(rc, "gjx_rmhmc_monge_funnel_step")
        else:
            params = logistic_params(logdensity_fn, C, L=L, step_size=step_size, max_iters=solver["max_iters"],
                                     convergence_tol=solver["convergence_tol"],
                                     divergence_tol=solver["divergence_tol"],
                                     divergence_threshold=divergence_threshold, seed=seed, counter=counter,
                                     dtype=position.dtype,
                                     waves_per_chain=_waves(waves_per_chain, logdensity_fn, C, position.dtype,
                                                            step_size, L, solver["max_iters"]))
            data = logdensity_fn.data(dtype=position.dtype, num_chains=C)
            rc = _lib.load().gjx_rmhmc_fisher_logistic_step(params, data, io, _lib.stream_handle(dev))
            _lib.check(rc, "gjx_rmhmc_fisher_logistic_step")

        def sq(t):
            return t[0] if single else t

        new_state = RMHMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]))
        pstate = IntegratorState(sq(out["proposal_position"]), sq(out["proposal_momentum"]),
                                 sq(out["proposal_velocity"]), sq(out["proposal_logdensity"]),
                                 sq(out["proposal_logdensity_grad"]))
        proposal = Proposal(pstate, sq(out["energy"]), sq(out["proposal_weight"]),
                            sq(out["proposal_sum_log_p_accept"]))
        info = RMHMCInfo(sq(out["momentum"]), sq(out["acceptance_rate"]), sq(out["is_accepted"].bool()),
                         sq(out["is_divergent"].bool()), sq(out["energy"]), proposal,
                         L if steps is None else sq(steps))
        if return_fixed_point_info:
            fp = FixedPointIterationInfo(sq(out["fp_success"][:, :L].bool()), sq(out["fp_norm"][:, :L]),
                                         sq(out["fp_iters"][:, :L]))
            return new_state, info, fp
        return new_state, info

    kernel.solver_kwargs = solver
    kernel.divergence_threshold = divergence_threshold
    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples, *,
           integrator=implicit_midpoint, divergence_threshold=1000, store_samples=True, lanes_per_chain=0,
           waves_per_chain=0, samples_out=None, max_num_integration_steps=None):
    """``num_samples`` transitions of every chain in ONE persistent launch.

    The reference runs this loop as ``jax.lax.scan`` over ``jax.vmap(kernel)``
    (examples/funnel/main.py:16-25); here the chain state stays in registers
    across transitions.  ``rng_key`` is a PRNGKey (device Philox; transition t
    uses counter ``key.counter + t``) or HostNoise with z [T, C, D], u [T, C].
    ``samples_out`` (a contiguous [T, C, D] tensor of the state dtype) receives
    the samples in place of a new allocation.  ``num_integration_steps`` is an
    int or, for the dynamic kernel, an integer tensor [T, C] of per-transition,
    per-chain counts.  Returns ``SampleResult``.
    """
    from .lmcmc_dense import is_dense_metric_fn
    from .sampling import SampleResult

    if is_dense_metric_fn(logdensity_fn, metric_fn):
        from . import rmhmc_dense

        return rmhmc_dense.sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps,
                                  num_samples, _solver_kwargs(integrator), divergence_threshold=divergence_threshold,
                                  store_samples=store_samples, samples_out=samples_out,
                                  max_num_integration_steps=max_num_integration_steps)

    solver = _solver_kwargs(integrator)
    position, single = _as_batch(state.position)
    if single:
        raise ValueError("sample() runs a batch of chains: position must be [C, D]")
    kind = _problem(logdensity_fn, metric_fn, position)
    position = position.contiguous()
    dev = position.device
    C, D = position.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
    L, steps = _steps(num_integration_steps, (T, C), dev, max_num_integration_steps)
    if isinstance(rng_key, HostNoise):
        z = rng_key.z.reshape(T, C, D).to(device=dev, dtype=position.dtype).contiguous()
        u = rng_key.u.reshape(T, C).to(device=dev, dtype=torch.float64).contiguous()
        seed, counter = 0, 0
    elif isinstance(rng_key, PRNGKey):
        z = u = None
        seed, counter = rng_key.seed, rng_key.counter
    else:
        raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")
    final_q = torch.empty_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.empty_like(position)
    if samples_out is not None:
        if samples_out.shape != (T, C, D) or samples_out.dtype != position.dtype or not samples_out.is_contiguous():
            raise ValueError(f"samples_out must be a contiguous [{T}, {C}, {D}] {position.dtype} tensor")
        samples = samples_out
    else:
        samples = torch.empty(T, C, D, dtype=position.dtype, device=dev) if store_samples else None
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.empty(C, dtype=torch.int32, device=dev)
    n_fev = torch.empty(C, dtype=torch.int64, device=dev)
    io = _lib.RmhmcSampleIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(), z=_lib.ptr(z),
                            u=_lib.ptr(u), final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                            final_logdensity_grad=final_g.data_ptr(), samples=_lib.ptr(samples),
                            acceptance_rate=acc.data_ptr(), num_accepted=n_acc.data_ptr(),
                            num_fevals=n_fev.data_ptr(), num_integration_steps=_lib.ptr(steps))
    if kind == "funnel_monge":
        params = _params(C, D, _dtype_code(position), step_size, L, solver, metric_fn.alpha2,
                         logdensity_fn.sigma, divergence_threshold, T=T, seed=seed,
                         lanes_per_chain=_lanes(lanes_per_chain, "rmhmc_monge_funnel", D, C, position.dtype),
                         counter=counter)
        rc = _lib.load().gjx_rmhmc_monge_funnel_sample(params, io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_rmhmc_monge_funnel_sample")
    else:
        params = logistic_params(logdensity_fn, C, L=L, T=T, step_size=step_size,
                                 max_iters=solver["max_iters"], convergence_tol=solver["convergence_tol"],
                                 divergence_tol=solver["divergence_tol"], divergence_threshold=divergence_threshold,
                                 seed=seed, counter=counter, dtype=position.dtype,
                                 waves_per_chain=_waves(waves_per_chain, logdensity_fn, C, position.dtype,
                                                            step_size, L, solver["max_iters"]))
        data = logdensity_fn.data(dtype=position.dtype, num_chains=C)
        rc = _lib.load().gjx_rmhmc_fisher_logistic_sample(params, data, io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_rmhmc_fisher_logistic_sample")
    return SampleResult(RMHMCState(final_q, final_l, final_g), samples, acc, n_acc, n_fev)


class rmhmc:
    """geomjax/rmhmc/rmhmc.py:247-311.

    The returned ``SamplingAlgorithm`` also carries ``sample_fn(rng_key,
    state, num_samples, **kw)``, the persistent multi-transition launch used
    by ``geomjax_amd.sampling.run_chains``.
    """

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, num_integration_steps, *,
                divergence_threshold: int = 1000, integrator=implicit_midpoint) -> SamplingAlgorithm:
        from .sampling import BoundAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps)

        def sample_fn(rng_key, state, num_samples, **kw):
            return sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples,
                          integrator=integrator, divergence_threshold=divergence_threshold, **kw)

        return BoundAlgorithm(init_fn, step_fn, sample_fn)


# ---------------------------------------------------------------------------
# Dynamic RMHMC (geomjax/rmhmc/rmhmc.py:179-244, 314-373)
# ---------------------------------------------------------------------------


class DynamicRMHMCState(NamedTuple):  # geomjax/rmhmc/rmhmc.py:44-55 (same fields and order)
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    random_generator_arg: torch.Tensor  # [C] int64, one per chain


def _next_random_arg(arg):
    """Default successor of a chain's random argument (the reference splits a
    JAX key; here the argument is an int64 counter per chain)."""
    return arg + 1


def _integration_steps(arg):
    """Default number of steps: uniform on {1, ..., 9} (the reference's
    ``jax.random.randint(key, (), 1, 10)``), from a multiplicative hash of the
    chain's counter (Knuth's 2654435761, low 32 bits), so it needs no host
    round trip."""
    h = (arg.to(torch.int64) * 2654435761) & 0xFFFFFFFF
    return (1 + h % 9).to(torch.int32)


def init_dynamic(position, logdensity_fn, random_generator_arg):
    """geomjax/rmhmc/rmhmc.py:99-106; ``random_generator_arg`` is a scalar or a [C]
    int64 tensor (give each chain its own, e.g. ``torch.arange(C)``)."""
    s = init(position, logdensity_fn)
    arg = torch.as_tensor(random_generator_arg, dtype=torch.int64, device=s.position.device)
    C = s.position.shape[0] if s.position.dim() == 2 else 1
    return DynamicRMHMCState(s.position, s.logdensity, s.logdensity_grad, arg.expand(C).clone())


def build_dynamic_kernel(integrator=implicit_midpoint, divergence_threshold: float = 1000,
                         next_random_arg_fn=_next_random_arg, integration_steps_fn=_integration_steps,
                         max_num_integration_steps=9, **kernel_kwargs):
    """geomjax/rmhmc/rmhmc.py:179-244: the RMHMC kernel with a number of
    integration steps drawn per chain from its ``random_generator_arg``.

    ``integration_steps_fn(arg, **kw)`` maps the [C] int64 arguments to [C]
    step counts (a torch function on the device); ``max_num_integration_steps``
    bounds them (values above are clamped by the kernel; None reads the
    maximum back from the device)."""
    base = build_kernel(integrator, divergence_threshold, **kernel_kwargs)

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, **integration_steps_kwargs):
        steps = integration_steps_fn(state.random_generator_arg, **integration_steps_kwargs)
        new, info = base(rng_key, RMHMCState(state.position, state.logdensity, state.logdensity_grad),
                         logdensity_fn, step_size, metric_fn, steps,
                         max_num_integration_steps=max_num_integration_steps)
        return (DynamicRMHMCState(new.position, new.logdensity, new.logdensity_grad,
                                  next_random_arg_fn(state.random_generator_arg)), info)

    return kernel


class dynamic_rmhmc:
    """geomjax/rmhmc/rmhmc.py:314-373."""

    init = staticmethod(init_dynamic)
    build_kernel = staticmethod(build_dynamic_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, *, divergence_threshold: int = 1000,
                integrator=implicit_midpoint, next_random_arg_fn=_next_random_arg,
                integration_steps_fn=_integration_steps, max_num_integration_steps=9):
        from .sampling import BoundAlgorithm

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
            r = sample(rng_key, RMHMCState(state.position, state.logdensity, state.logdensity_grad),
                       logdensity_fn, step_size, metric_fn, torch.stack(args), num_samples, integrator=integrator,
                       divergence_threshold=divergence_threshold,
                       max_num_integration_steps=max_num_integration_steps, **kw)
            return r._replace(state=DynamicRMHMCState(r.state.position, r.state.logdensity,
                                                      r.state.logdensity_grad, a))

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
