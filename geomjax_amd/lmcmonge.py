"""Lagrangian Monte Carlo with the Monge metric, on the HIP kernels.

Mirrors geomjax/lmcmonge/lmc.py:32-405 and geomjax/lmcmonge/integrators.py:
``LMCState``, ``LMCInfo``, ``RiemannianIntegratorState``, ``init``,
``build_kernel`` returning ``kernel(rng_key, state, logdensity_fn,
step_size, inverse_mass_matrix, num_integration_steps, alpha2)`` and the
``lmc`` algorithm class (exported as ``geomjax_amd.lmcmonge`` like the
reference's alias, geomjax/__init__.py:24).

``dynamic_lmc`` (``DynamicLMCState``, ``init_dynamic``,
``build_dynamic_kernel``) is the reference's geomjax/lmcmonge/lmc.py:47-61,
112-127,240-309,408-469 with its state handling fixed: as written there the
kernel builds ``LMCState`` with an extra alpha2 field and never passes alpha2 to
the base kernel, so it cannot run; here the state's alpha2 is passed through
and kept.  Step counts are drawn per chain on the device (as
``geomjax_amd.rmhmc.dynamic_rmhmc``).

``reference_faithful=True`` (default) reproduces the reference bit for bit in
algebra, including the Sherman-Morrison correction without alpha2
(integrators.py:186-189) and the momentum's extra factor L (lmc.py:202-204);
``False`` uses the exact algebra for both.  Batching, descriptors and keys
are as in ``geomjax_amd.rmhmc``.
"""
from typing import NamedTuple

import torch

from . import _lib
from .metrics import _check_alpha2
from .models import NealFunnel, _dtype_code
from .proposal import Proposal
from .random import HostNoise, PRNGKey
from .rmhmc import _lanes

__all__ = ["LMCState", "LMCInfo", "RiemannianIntegratorState", "init", "build_kernel", "lmc", "lan_integrator",
           "sample", "DynamicLMCState", "init_dynamic", "build_dynamic_kernel", "dynamic_lmc",
           "proposal_alpha2_derivative"]


class RiemannianIntegratorState(NamedTuple):  # geomjax/lmcmonge/integrators.py:26-44
    alpha2: float
    position: torch.Tensor
    momentum: torch.Tensor
    velocity: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad_norm: torch.Tensor
    dl_ig: torch.Tensor
    Hdl_ig: torch.Tensor
    ig_Hdl_ig: torch.Tensor
    logdensity_hvp_velocity_norm: torch.Tensor
    determinant_metric: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCState(NamedTuple):  # geomjax/lmcmonge/lmc.py:32-44
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor


class LMCInfo(NamedTuple):  # geomjax/lmcmonge/lmc.py:63-98
    velocity: torch.Tensor
    acceptance_rate: torch.Tensor
    is_accepted: torch.Tensor
    is_divergent: torch.Tensor
    energy: torch.Tensor
    proposal: Proposal
    num_integration_steps: int


def lan_integrator(*args, **kwargs):
    """Marker for Lan's explicit Lagrangian integrator (integrators.py:52-155); runs inside the HIP kernel."""
    raise TypeError("lan_integrator is executed inside the HIP kernel; pass it to build_kernel/lmc")


def init(position, logdensity_fn):
    """geomjax/lmcmonge/lmc.py:101-109."""
    if not isinstance(logdensity_fn, NealFunnel):
        raise NotImplementedError(f"no HIP log density for {logdensity_fn!r}")
    logdensity, grad = logdensity_fn.value_and_grad(position)
    return LMCState(position, logdensity, grad, torch.zeros_like(logdensity))


def _check(logdensity_fn, position, inverse_mass_matrix):
    if not isinstance(logdensity_fn, NealFunnel):
        raise NotImplementedError(f"no HIP LMC-Monge kernel for logdensity_fn={logdensity_fn!r}")
    if position.shape[-1] != logdensity_fn.D:
        raise ValueError(f"position has {position.shape[-1]} coordinates, the model has D={logdensity_fn.D}")
    if not position.is_cuda:
        raise ValueError("LMC states must live on a HIP device (cuda:N)")
    im = torch.as_tensor(inverse_mass_matrix)
    if im.dim() != 1:
        # geomjax/lmcmonge/metrics.py:141-153 allows only a diagonal mass
        raise ValueError(f"The mass matrix has the wrong number of dimensions: expected 1, got {im.dim()}.")
    if im.shape[0] != position.shape[-1]:
        raise ValueError("inverse_mass_matrix must have one entry per coordinate")
    return im.to(device=position.device, dtype=position.dtype).contiguous()


def _params(C, D, dtype_code, step_size, L, alpha2, sigma, divergence_threshold, faithful, T=1, lpc=0, seed=0,
            counter=0):
    flags = 0 if faithful else (_lib.LMC_EXACT_SHERMAN_MORRISON | _lib.LMC_EXACT_MOMENTUM)
    return _lib.FunnelParams(
        num_chains=C, dim=D, dtype=dtype_code, num_integration_steps=int(L), max_iters=0, num_transitions=int(T),
        lanes_per_chain=int(lpc), flags=flags, step_size=float(step_size), alpha2=float(alpha2), sigma=float(sigma),
        convergence_tol=0.0, divergence_tol=0.0, divergence_threshold=float(divergence_threshold), seed=int(seed),
        counter=int(counter))


def _noise(rng_key, T, C, D, dev, dtype):
    if isinstance(rng_key, HostNoise):
        shape = (C, D) if T is None else (T, C, D)
        ushape = (C,) if T is None else (T, C)
        z = rng_key.z.reshape(shape).to(device=dev, dtype=dtype).contiguous()
        u = rng_key.u.reshape(ushape).to(device=dev, dtype=torch.float64).contiguous()
        return z, u, 0, 0
    if isinstance(rng_key, PRNGKey):
        return None, None, rng_key.seed, rng_key.counter
    raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")


def build_kernel(integrator=lan_integrator, divergence_threshold: float = 1000, *, reference_faithful=True,
                 lanes_per_chain=0):
    """geomjax/lmcmonge/lmc.py:130-237."""
    if integrator is not lan_integrator:
        raise NotImplementedError("only lan_integrator is implemented on the GPU")

    def kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps, alpha2, *,
               max_num_integration_steps=None):
        """``num_integration_steps``: an int, or a [C] integer tensor of per-chain
        counts (the dynamic kernel's form), bounded by ``max_num_integration_steps``."""
        from .rmhmc import _steps

        _check_alpha2(alpha2)
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        im = _check(logdensity_fn, position, inverse_mass_matrix)
        dev = position.device
        C, D = position.shape
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        grad = state.logdensity_grad.reshape(C, D).to(position.dtype).contiguous()
        J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
        z, u, seed, counter = _noise(rng_key, None, C, D, dev, position.dtype)
        L, steps = _steps(num_integration_steps, (C,), dev, max_num_integration_steps)

        def vec():
            return torch.empty(C, D, dtype=position.dtype, device=dev)

        def sca(dt=torch.float64):
            return torch.empty(C, dtype=dt, device=dev)

        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), new_volume_adjustment=sca(),
                   velocity=vec(), acceptance_rate=sca(), is_accepted=sca(torch.uint8),
                   is_divergent=sca(torch.uint8), energy=sca(), proposal_position=vec(), proposal_momentum=vec(),
                   proposal_velocity=vec(), proposal_logdensity=sca(), proposal_logdensity_grad_norm=vec(),
                   proposal_dl_ig=vec(), proposal_Hdl_ig=vec(), proposal_ig_Hdl_ig=vec(),
                   proposal_logdensity_hvp_velocity_norm=vec(), proposal_determinant_metric=sca(),
                   proposal_volume_adjustment=sca(), proposal_weight=sca(), proposal_sum_log_p_accept=sca())
        io = _lib.LmcStepIO(inverse_mass_matrix=im.data_ptr(), position=position.data_ptr(),
                            logdensity=logdensity.data_ptr(), logdensity_grad=grad.data_ptr(),
                            volume_adjustment=J.data_ptr(), z=_lib.ptr(z), u=_lib.ptr(u),
                            num_integration_steps=_lib.ptr(steps), **{k: v.data_ptr() for k, v in out.items()})
        params = _params(C, D, _dtype_code(position), step_size, L, alpha2, logdensity_fn.sigma,
                         divergence_threshold, reference_faithful, seed=seed, counter=counter,
                         lpc=_lanes(lanes_per_chain, "lmcmonge_funnel", D, C, position.dtype))
        rc = _lib.load().gjx_lmcmonge_funnel_step(params, io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_lmcmonge_funnel_step")

        def sq(t):
            return t[0] if single else t

        new_state = LMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]),
                             sq(out["new_volume_adjustment"]))
        pstate = RiemannianIntegratorState(
            float(alpha2), sq(out["proposal_position"]), sq(out["proposal_momentum"]), sq(out["proposal_velocity"]),
            sq(out["proposal_logdensity"]), sq(out["proposal_logdensity_grad_norm"]), sq(out["proposal_dl_ig"]),
            sq(out["proposal_Hdl_ig"]), sq(out["proposal_ig_Hdl_ig"]),
            sq(out["proposal_logdensity_hvp_velocity_norm"]), sq(out["proposal_determinant_metric"]),
            sq(out["proposal_volume_adjustment"]))
        proposal = Proposal(pstate, sq(out["energy"]), sq(out["proposal_weight"]),
                            sq(out["proposal_sum_log_p_accept"]))
        info = LMCInfo(sq(out["velocity"]), sq(out["acceptance_rate"]), sq(out["is_accepted"].bool()),
                       sq(out["is_divergent"].bool()), sq(out["energy"]), proposal,
                       L if steps is None else sq(steps))
        return new_state, info

    return kernel


def proposal_alpha2_derivative(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps,
                               alpha2, *, reference_faithful=True, lanes_per_chain=0, max_num_integration_steps=None):
    """(proposal position, d proposal position / d alpha2) of one LMC-Monge
    transition per chain: the jax.jacfwd of the proposal position in alpha2
    that geomjax/adaptation/chees_adaptation_lmcmonge.py:475-519 takes, in
    forward mode on the device (float64, alpha2 > 0).  ``rng_key`` as for the
    step kernel (only the normals of a ``HostNoise`` are used), so the result
    belongs to the trajectory the step kernel integrates with the same key."""
    from .rmhmc import _steps

    _check_alpha2(alpha2)
    position = state.position
    single = position.dim() == 1
    position = (position[None] if single else position).contiguous()
    if position.dtype != torch.float64:
        raise TypeError("the alpha2 derivative runs in float64")
    im = _check(logdensity_fn, position, inverse_mass_matrix)
    dev = position.device
    C, D = position.shape
    z, _, seed, counter = _noise(rng_key, None, C, D, dev, torch.float64)
    L, steps = _steps(num_integration_steps, (C,), dev, max_num_integration_steps)
    prop = torch.empty(C, D, dtype=torch.float64, device=dev)
    dprop = torch.empty(C, D, dtype=torch.float64, device=dev)
    io = _lib.LmcAlpha2TangentIO(inverse_mass_matrix=im.data_ptr(), position=position.data_ptr(), z=_lib.ptr(z),
                                 num_integration_steps=_lib.ptr(steps), proposal_position=prop.data_ptr(),
                                 d_proposal_position_d_alpha2=dprop.data_ptr())
    params = _params(C, D, _lib.GJX_F64, step_size, L, alpha2, logdensity_fn.sigma, 1000.0, reference_faithful,
                     lpc=lanes_per_chain, seed=seed, counter=counter)
    rc = _lib.load().gjx_lmcmonge_funnel_alpha2_tangent(params, io, _lib.stream_handle(dev))
    _lib.check(rc, "gjx_lmcmonge_funnel_alpha2_tangent")
    return (prop[0], dprop[0]) if single else (prop, dprop)


def sample(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps, alpha2,
           num_samples, *, divergence_threshold=1000, reference_faithful=True, store_samples=True, lanes_per_chain=0,
           samples_out=None, max_num_integration_steps=None):
    """``num_samples`` LMC-Monge transitions of every chain in one persistent launch
    (``num_integration_steps`` an int or a [T, C] integer tensor)."""
    from .rmhmc import _steps
    from .sampling import SampleResult

    _check_alpha2(alpha2)
    position = state.position
    if position.dim() != 2:
        raise ValueError("sample() runs a batch of chains: position must be [C, D]")
    position = position.contiguous()
    im = _check(logdensity_fn, position, inverse_mass_matrix)
    dev = position.device
    C, D = position.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
    J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
    z, u, seed, counter = _noise(rng_key, T, C, D, dev, position.dtype)
    L, steps = _steps(num_integration_steps, (T, C), dev, max_num_integration_steps)
    final_q = torch.empty_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.empty_like(position)
    final_J = torch.empty(C, dtype=torch.float64, device=dev)
    if samples_out is not None:
        if samples_out.shape != (T, C, D) or samples_out.dtype != position.dtype or not samples_out.is_contiguous():
            raise ValueError(f"samples_out must be a contiguous [{T}, {C}, {D}] {position.dtype} tensor")
        samples = samples_out
    else:
        samples = torch.empty(T, C, D, dtype=position.dtype, device=dev) if store_samples else None
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.empty(C, dtype=torch.int32, device=dev)
    io = _lib.LmcSampleIO(inverse_mass_matrix=im.data_ptr(), position=position.data_ptr(),
                          logdensity=logdensity.data_ptr(), volume_adjustment=J.data_ptr(), z=_lib.ptr(z),
                          u=_lib.ptr(u), final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                          final_logdensity_grad=final_g.data_ptr(), final_volume_adjustment=final_J.data_ptr(),
                          samples=_lib.ptr(samples), acceptance_rate=acc.data_ptr(), num_accepted=n_acc.data_ptr(),
                          num_integration_steps=_lib.ptr(steps))
    params = _params(C, D, _dtype_code(position), step_size, L, alpha2, logdensity_fn.sigma,
                     divergence_threshold, reference_faithful, T=T, seed=seed, counter=counter,
                     lpc=_lanes(lanes_per_chain, "lmcmonge_funnel", D, C, position.dtype))
    rc = _lib.load().gjx_lmcmonge_funnel_sample(params, io, _lib.stream_handle(dev))
    _lib.check(rc, "gjx_lmcmonge_funnel_sample")
    # every LMC step evaluates the vector field once (one gradient + Hessian-vector products)
    if steps is None:
        fev = torch.full((C,), T * L, dtype=torch.int64, device=dev)
    else:
        fev = steps.clamp(0, L).to(torch.int64).sum(0)
    return SampleResult(LMCState(final_q, final_l, final_g, final_J), samples, acc, n_acc, fev)


class lmc:
    """geomjax/lmcmonge/lmc.py:312-405 (exported as ``lmcmonge``)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps, *, alpha2=0.001,
                divergence_threshold: int = 1000, integrator=lan_integrator, reference_faithful=True):
        from .sampling import BoundAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold, reference_faithful=reference_faithful)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps,
                          alpha2)

        def sample_fn(rng_key, state, num_samples, **kw):
            return sample(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, num_integration_steps,
                          alpha2, num_samples, divergence_threshold=divergence_threshold,
                          reference_faithful=reference_faithful, **kw)

        return BoundAlgorithm(init_fn, step_fn, sample_fn)


# ---------------------------------------------------------------------------
# Dynamic LMC-Monge (geomjax/lmcmonge/lmc.py:47-61,112-127,240-309,408-469)
# ---------------------------------------------------------------------------


class DynamicLMCState(NamedTuple):  # geomjax/lmcmonge/lmc.py:47-61 (same fields and order)
    alpha2: float
    position: torch.Tensor
    logdensity: torch.Tensor
    logdensity_grad: torch.Tensor
    volume_adjustment: torch.Tensor
    random_generator_arg: torch.Tensor  # [C] int64, one per chain


def init_dynamic(position, logdensity_fn, alpha2, random_generator_arg):
    """geomjax/lmcmonge/lmc.py:112-127; ``random_generator_arg`` is a scalar or a
    [C] int64 tensor."""
    _check_alpha2(alpha2)
    s = init(position, logdensity_fn)
    arg = torch.as_tensor(random_generator_arg, dtype=torch.int64, device=s.position.device)
    C = s.position.shape[0] if s.position.dim() == 2 else 1
    return DynamicLMCState(float(alpha2), s.position, s.logdensity, s.logdensity_grad, s.volume_adjustment,
                           arg.expand(C).clone())


def build_dynamic_kernel(integrator=lan_integrator, divergence_threshold: float = 1000, next_random_arg_fn=None,
                         integration_steps_fn=None, max_num_integration_steps=9, **kernel_kwargs):
    """geomjax/lmcmonge/lmc.py:240-309: the LMC-Monge kernel with a number of
    integration steps drawn per chain from its ``random_generator_arg``, run
    at the state's alpha2 (which the new state keeps)."""
    from .rmhmc import _integration_steps, _next_random_arg

    next_random_arg_fn = next_random_arg_fn or _next_random_arg
    integration_steps_fn = integration_steps_fn or _integration_steps
    base = build_kernel(integrator, divergence_threshold, **kernel_kwargs)

    def kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, **integration_steps_kwargs):
        steps = integration_steps_fn(state.random_generator_arg, **integration_steps_kwargs)
        lmc_state = LMCState(state.position, state.logdensity, state.logdensity_grad, state.volume_adjustment)
        new, info = base(rng_key, lmc_state, logdensity_fn, step_size, inverse_mass_matrix, steps, state.alpha2,
                         max_num_integration_steps=max_num_integration_steps)
        return (DynamicLMCState(state.alpha2, new.position, new.logdensity, new.logdensity_grad,
                                new.volume_adjustment, next_random_arg_fn(state.random_generator_arg)), info)

    return kernel


class dynamic_lmc:
    """geomjax/lmcmonge/lmc.py:408-469."""

    init = staticmethod(init_dynamic)
    build_kernel = staticmethod(build_dynamic_kernel)

    def __new__(cls, logdensity_fn, step_size, inverse_mass_matrix, *, alpha2=0.001, divergence_threshold: int = 1000,
                integrator=lan_integrator, next_random_arg_fn=None, integration_steps_fn=None,
                max_num_integration_steps=9, reference_faithful=True):
        from .rmhmc import _integration_steps, _next_random_arg
        from .sampling import BoundAlgorithm

        next_random_arg_fn = next_random_arg_fn or _next_random_arg
        integration_steps_fn = integration_steps_fn or _integration_steps
        kernel = cls.build_kernel(integrator, divergence_threshold, next_random_arg_fn, integration_steps_fn,
                                  max_num_integration_steps, reference_faithful=reference_faithful)

        def init_fn(position, random_generator_arg):
            return cls.init(position, logdensity_fn, alpha2, random_generator_arg)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix)

        def sample_fn(rng_key, state, num_samples, **kw):
            """``num_samples`` transitions in one persistent launch; the step counts
            of all transitions are drawn up front from the successive arguments."""
            args, a = [], state.random_generator_arg
            for _ in range(int(num_samples)):
                args.append(integration_steps_fn(a))
                a = next_random_arg_fn(a)
            r = sample(rng_key, LMCState(state.position, state.logdensity, state.logdensity_grad,
                                         state.volume_adjustment), logdensity_fn, step_size, inverse_mass_matrix,
                       torch.stack(args), state.alpha2, num_samples, divergence_threshold=divergence_threshold,
                       reference_faithful=reference_faithful, max_num_integration_steps=max_num_integration_steps,
                       **kw)
            return r._replace(state=DynamicLMCState(state.alpha2, r.state.position, r.state.logdensity,
                                                    r.state.logdensity_grad, r.state.volume_adjustment, a))

        return BoundAlgorithm(init_fn, step_fn, sample_fn)
