"""NUTS with the RMHMC implicit-midpoint integrator (geomjax/rmhmc/nuts.py).

Same API as the reference: ``nuts(logdensity_fn, step_size, metric_fn, *,
max_num_doublings=10, divergence_threshold=1000, integrator=implicit_midpoint,
stopping_criterion="euc")`` with ``init`` (= rmhmc.init) and ``step``; the
kernel ``build_kernel(integrator, divergence_threshold)(rng_key, state,
logdensity_fn, step_size, metric_fn, max_num_doublings, stopping_criterion)``.
The whole transition (momentum draw, iterative tree doubling, progressive
multinomial sampling, U-turn checks) runs in one HIP kernel, float64 for a
float64 state and the tuned form for a float32 one (float32 state and
trajectory, float64 energies and tree weights), with the three U-turn
criteria of geomjax/rmhmc/metrics.py:100-117 ("euc", the default; "bet";
"riem").  Served: Neal's funnel with the Monge metric (one wave per chain) and
Bayesian logistic regression with the Fisher metric, one posterior or a
``LogisticBatch``, every posterior with D <= 64 (float64: one 256-lane
workgroup per chain; tuned: one wave per chain, split-bf16 MFMA passes).

``rng_key`` is a ``PRNGKey`` (device Philox) or ``NUTSNoise(z, u_direction,
u_merge, u_step)`` with the draws the reference takes from its keys: the
momentum normals [C, D], per expansion e the direction uniform (+1 when
< 0.5) and the biased-merge uniform [C, M], and per integration step j of the
transition the progressive-sampling uniform [C, 2^M - 1].

``NUTSInfo.trajectory_leftmost_state`` / ``trajectory_rightmost_state`` carry
the positions of the trajectory's two ends (the reference keeps the whole
IntegratorState there).
"""
import functools
from typing import NamedTuple

import torch

from . import _lib
from .metrics import FisherMetric, MongeMetric
from .models import LogisticBatch, LogisticRegression, NealFunnel, logistic_params
from .random import PRNGKey
from .rmhmc import RMHMCState, _params, _solver_kwargs, implicit_midpoint, init


class NUTSNoise(NamedTuple):
    z: torch.Tensor            # [C, D]
    u_direction: torch.Tensor  # [C, M]
    u_merge: torch.Tensor      # [C, M]
    u_step: torch.Tensor       # [C, 2^M - 1]


class NUTSSampleResult(NamedTuple):
    """SampleResult of a persistent NUTS launch, with the trajectory lengths."""
    state: RMHMCState
    samples: torch.Tensor                # [T, C, D] or None
    acceptance_rate: torch.Tensor        # [T, C] float32
    num_accepted: None                   # NUTS samples from the trajectory: no accept count
    num_fevals: torch.Tensor             # [C] int64: integration steps of the launch
    num_integration_steps: torch.Tensor  # [T, C] int32


class NUTSInfo(NamedTuple):  # geomjax/rmhmc/nuts.py:36-75 (same fields and order)
    momentum: torch.Tensor
    is_divergent: torch.Tensor
    is_turning: torch.Tensor
    energy: torch.Tensor
    trajectory_leftmost_state: torch.Tensor
    trajectory_rightmost_state: torch.Tensor
    num_trajectory_expansions: torch.Tensor
    num_integration_steps: torch.Tensor
    acceptance_rate: torch.Tensor


def build_kernel(integrator=implicit_midpoint, divergence_threshold: int = 1000):
    """geomjax/rmhmc/nuts.py:78-161.  Plain callables (a user log density and
    dense metric) run the batched dense-metric NUTS of ``nuts_dense``."""
    solver = _solver_kwargs(integrator)

    def _dense_kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings, stopping_criterion):
        from . import nuts_dense
        from .rmhmc import RMHMCState

        single = state.position.dim() == 1
        st = RMHMCState(*((t[None] if single else t) for t in state)) if single else state
        C, D = st.position.shape
        dev = st.position.device
        M = int(max_num_doublings)
        if isinstance(rng_key, NUTSNoise):
            f = lambda t, *shape: t.reshape(*shape).to(device=dev, dtype=torch.float64).contiguous()
            noise = (f(rng_key.z, C, D), f(rng_key.u_direction, C, M), f(rng_key.u_merge, C, M),
                     f(rng_key.u_step, C, 2 ** M - 1))
        elif isinstance(rng_key, PRNGKey):
            gen = torch.Generator(device=dev)
            gen.manual_seed((int(rng_key.seed) * 0x9E3779B1 + int(rng_key.counter)) & 0x7FFFFFFFFFFFFFFF)
            noise = (torch.randn(C, D, generator=gen, dtype=torch.float64, device=dev),
                     *(torch.rand(C, n, generator=gen, dtype=torch.float64, device=dev) for n in (M, M, 2 ** M - 1)))
        else:
            raise TypeError("rng_key must be NUTSNoise or PRNGKey")
        q, l, g, inf = nuts_dense.transition(noise, st, logdensity_fn, step_size, metric_fn, solver, M,
                                             stopping_criterion, divergence_threshold)
        sq = (lambda t: t[0]) if single else (lambda t: t)
        info = NUTSInfo(sq(inf["momentum"]), sq(inf["is_divergent"]), sq(inf["is_turning"]), sq(inf["energy"]),
                        sq(inf["left"]), sq(inf["right"]), sq(inf["num_trajectory_expansions"]),
                        sq(inf["num_integration_steps"]), sq(inf["acceptance_rate"]))
        return RMHMCState(sq(q), sq(l), sq(g)), info

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings: int = 10,
               stopping_criterion: str = "euc"):
        if stopping_criterion not in _lib.GJX_UTURN:
            raise ValueError(f"stopping_criterion must be one of {sorted(_lib.GJX_UTURN)} (got {stopping_criterion!r})")
        from .lmcmc_dense import is_dense_metric_fn

        if is_dense_metric_fn(logdensity_fn, metric_fn):
            return _dense_kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings,
                                 stopping_criterion)
        funnel = isinstance(logdensity_fn, NealFunnel) and isinstance(metric_fn, MongeMetric)
        logistic = isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) and isinstance(metric_fn, FisherMetric)
        if not (funnel or logistic):
            raise NotImplementedError("the NUTS kernels serve Neal's funnel with the Monge metric and logistic "
                                      "regression with the Fisher metric")
        position = state.position
        single = position.dim() == 1
        if single:
            position = position[None]
        position = position.contiguous()
        if position.dtype not in (torch.float64, torch.float32) or not position.is_cuda:
            raise ValueError("NUTS positions must be float64 (plain) or float32 (tuned) tensors on a HIP device")
        C, D = position.shape
        if D != logdensity_fn.D:
            raise ValueError(f"position has {D} coordinates, the model has {logdensity_fn.D}")
        if isinstance(logdensity_fn, LogisticBatch) and C != logdensity_fn.num_chains:
            raise ValueError(f"the batch has {logdensity_fn.num_chains} chains, got {C}")
        dev = position.device
        M = int(max_num_doublings)
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        if isinstance(rng_key, NUTSNoise):
            f = lambda t, *shape: t.reshape(*shape).to(device=dev, dtype=torch.float64).contiguous()
            noise = (rng_key.z.reshape(C, D).to(device=dev, dtype=position.dtype).contiguous(),
                     f(rng_key.u_direction, C, M), f(rng_key.u_merge, C, M), f(rng_key.u_step, C, 2 ** M - 1))
            seed, counter = 0, 0
        elif isinstance(rng_key, PRNGKey):
            noise = (None, None, None, None)
            seed, counter = rng_key.seed, rng_key.counter
        else:
            raise TypeError("rng_key must be NUTSNoise or PRNGKey")
        vec = lambda: torch.empty(C, D, dtype=position.dtype, device=dev)
        sca = lambda dt=torch.float64: torch.empty(C, dtype=dt, device=dev)
        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), momentum=vec(),
                   energy=sca(), acceptance_rate=sca(), is_divergent=sca(torch.uint8), is_turning=sca(torch.uint8),
                   num_trajectory_expansions=sca(torch.int32), num_integration_steps=sca(torch.int32),
                   leftmost_position=vec(), rightmost_position=vec())
        io = _lib.NutsIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(), z=_lib.ptr(noise[0]),
                         u_direction=_lib.ptr(noise[1]), u_merge=_lib.ptr(noise[2]), u_step=_lib.ptr(noise[3]),
                         **{k: v.data_ptr() for k, v in out.items()})
        crit = _lib.GJX_UTURN[stopping_criterion]
        if funnel:
            code = _lib.GJX_F64 if position.dtype == torch.float64 else _lib.GJX_F32
            params = _params(C, D, code, step_size, 0, solver, metric_fn.alpha2, logdensity_fn.sigma,
                             divergence_threshold, seed=seed, counter=counter)
            rc = _lib.load().gjx_rmhmc_monge_funnel_nuts(params, M, crit, io, _lib.stream_handle(dev))
            _lib.check(rc, "gjx_rmhmc_monge_funnel_nuts")
        else:
            params = logistic_params(logdensity_fn, C, step_size=step_size, max_iters=solver["max_iters"],
                                     convergence_tol=solver["convergence_tol"],
                                     divergence_tol=solver["divergence_tol"],
                                     divergence_threshold=divergence_threshold, seed=seed, counter=counter,
                                     dtype=position.dtype)
            data = logdensity_fn.data(dtype=position.dtype, num_chains=C)
            rc = _lib.load().gjx_rmhmc_fisher_logistic_nuts(params, data, M, crit, io, _lib.stream_handle(dev))
            _lib.check(rc, "gjx_rmhmc_fisher_logistic_nuts")
        sq = (lambda t: t[0]) if single else (lambda t: t)
        new_state = RMHMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]))
        info = NUTSInfo(sq(out["momentum"]), sq(out["is_divergent"].bool()), sq(out["is_turning"].bool()),
                        sq(out["energy"]), sq(out["leftmost_position"]), sq(out["rightmost_position"]),
                        sq(out["num_trajectory_expansions"]), sq(out["num_integration_steps"]),
                        sq(out["acceptance_rate"]))
        return new_state, info

    return kernel


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples, *, max_num_doublings=10,
           stopping_criterion="euc", integrator=implicit_midpoint, divergence_threshold=1000, store_samples=True,
           samples_out=None):
    """``num_samples`` NUTS transitions of every chain in ONE launch (Neal's
    funnel with the Monge metric): the reference's loop over the kernel
    (examples/funnel/main.py:16-25).  The chains keep their state on chip, so
    the launch lasts as long as the largest per-chain sum of trajectory
    lengths.  ``rng_key``: a PRNGKey (transition t draws at counter + t) or
    NUTSNoise with a leading [T] axis on every array.  A float32 state runs
    the tuned form (float32 state and trajectory, float64 energies).  Returns
    ``SampleResult`` (num_fevals = integration steps per chain)."""
    if not (isinstance(logdensity_fn, NealFunnel) and isinstance(metric_fn, MongeMetric)):
        raise NotImplementedError("the persistent NUTS launch serves Neal's funnel with the Monge metric")
    if stopping_criterion not in _lib.GJX_UTURN:
        raise ValueError(f"stopping_criterion must be one of {sorted(_lib.GJX_UTURN)}")
    solver = _solver_kwargs(integrator)
    position = state.position.contiguous()
    if position.dim() != 2 or position.dtype not in (torch.float64, torch.float32) or not position.is_cuda:
        raise ValueError("sample() runs a batch of chains: position must be a float64 or float32 [C, D] tensor "
                         "on a HIP device")
    C, D = position.shape
    if D != logdensity_fn.D:
        raise ValueError(f"position has {D} coordinates, the model has {logdensity_fn.D}")
    T, M = int(num_samples), int(max_num_doublings)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    dev = position.device
    logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
    if isinstance(rng_key, NUTSNoise):
        f = lambda t, *shape: t.reshape(*shape).to(device=dev, dtype=torch.float64).contiguous()
        noise = (rng_key.z.reshape(T, C, D).to(device=dev, dtype=position.dtype).contiguous(),
                 f(rng_key.u_direction, T, C, M), f(rng_key.u_merge, T, C, M), f(rng_key.u_step, T, C, 2 ** M - 1))
        seed, counter = 0, 0
    elif isinstance(rng_key, PRNGKey):
        noise = (None, None, None, None)
        seed, counter = rng_key.seed, rng_key.counter
    else:
        raise TypeError("rng_key must be NUTSNoise or PRNGKey")
    if samples_out is not None:
        if samples_out.shape != (T, C, D) or samples_out.dtype != position.dtype or not samples_out.is_contiguous():
            raise ValueError(f"samples_out must be a contiguous [{T}, {C}, {D}] {position.dtype} tensor")
        samples = samples_out
    else:
        samples = torch.empty(T, C, D, dtype=position.dtype, device=dev) if store_samples else None
    final_q = torch.empty_like(position)
    final_l = torch.empty(C, dtype=torch.float64, device=dev)
    final_g = torch.empty_like(position)
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    steps = torch.empty(T, C, dtype=torch.int32, device=dev)
    fev = torch.empty(C, dtype=torch.int64, device=dev)
    io = _lib.NutsSampleIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(), z=_lib.ptr(noise[0]),
                           u_direction=_lib.ptr(noise[1]), u_merge=_lib.ptr(noise[2]), u_step=_lib.ptr(noise[3]),
                           final_position=final_q.data_ptr(), final_logdensity=final_l.data_ptr(),
                           final_logdensity_grad=final_g.data_ptr(), samples=_lib.ptr(samples),
                           acceptance_rate=acc.data_ptr(), num_integration_steps=steps.data_ptr(),
                           num_fevals=fev.data_ptr())
    code = _lib.GJX_F64 if position.dtype == torch.float64 else _lib.GJX_F32
    params = _params(C, D, code, step_size, 0, solver, metric_fn.alpha2, logdensity_fn.sigma,
                     divergence_threshold, T=T, seed=seed, counter=counter)
    rc = _lib.load().gjx_rmhmc_monge_funnel_nuts_sample(params, M, _lib.GJX_UTURN[stopping_criterion], io,
                                                         _lib.stream_handle(dev))
    _lib.check(rc, "gjx_rmhmc_monge_funnel_nuts_sample")
    return NUTSSampleResult(RMHMCState(final_q, final_l, final_g), samples, acc, None, fev, steps)


class nuts:
    """geomjax/rmhmc/nuts.py:164-250."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, *, max_num_doublings: int = 10,
                divergence_threshold: int = 1000, integrator=implicit_midpoint, stopping_criterion: str = "euc"):
        from .base import SamplingAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings,
                          stopping_criterion=stopping_criterion)

        if isinstance(logdensity_fn, NealFunnel) and isinstance(metric_fn, MongeMetric):
            from .sampling import BoundAlgorithm

            def sample_fn(rng_key, state, num_samples, **kw):
                return sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_samples,
                              max_num_doublings=max_num_doublings, stopping_criterion=stopping_criterion,
                              integrator=integrator, divergence_threshold=divergence_threshold, **kw)

            return BoundAlgorithm(init_fn, step_fn, sample_fn)
        return SamplingAlgorithm(init_fn, step_fn)


__all__ = ["nuts", "build_kernel", "sample", "NUTSInfo", "NUTSNoise", "init"]
