"""NUTS with the RMHMC implicit-midpoint integrator (geomjax/rmhmc/nuts.py).

Same API as the reference: ``nuts(logdensity_fn, step_size, metric_fn, *,
max_num_doublings=10, divergence_threshold=1000, integrator=implicit_midpoint,
stopping_criterion="euc")`` with ``init`` (= rmhmc.init) and ``step``; the
kernel ``build_kernel(integrator, divergence_threshold)(rng_key, state,
logdensity_fn, step_size, metric_fn, max_num_doublings, stopping_criterion)``.
The whole transition (momentum draw, iterative tree doubling, progressive
multinomial sampling, U-turn checks) runs in one HIP kernel, float64, with the
three U-turn criteria of geomjax/rmhmc/metrics.py:100-117 ("euc", the
default; "bet"; "riem").  Served: Neal's funnel with the Monge metric (one
wave per chain) and Bayesian logistic regression with the Fisher metric, one
posterior or a ``LogisticBatch``, every posterior with D <= 64 (one
workgroup per chain).

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
        if position.dtype != torch.float64 or not position.is_cuda:
            raise ValueError("NUTS positions must be float64 tensors on a HIP device")
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
            noise = (f(rng_key.z, C, D), f(rng_key.u_direction, C, M), f(rng_key.u_merge, C, M),
                     f(rng_key.u_step, C, 2 ** M - 1))
            seed, counter = 0, 0
        elif isinstance(rng_key, PRNGKey):
            noise = (None, None, None, None)
            seed, counter = rng_key.seed, rng_key.counter
        else:
            raise TypeError("rng_key must be NUTSNoise or PRNGKey")
        vec = lambda: torch.empty(C, D, dtype=torch.float64, device=dev)
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
            params = _params(C, D, _lib.GJX_F64, step_size, 0, solver, metric_fn.alpha2, logdensity_fn.sigma,
                             divergence_threshold, seed=seed, counter=counter)
            rc = _lib.load().gjx_rmhmc_monge_funnel_nuts(params, M, crit, io, _lib.stream_handle(dev))
            _lib.check(rc, "gjx_rmhmc_monge_funnel_nuts")
        else:
            params = logistic_params(logdensity_fn, C, step_size=step_size, max_iters=solver["max_iters"],
                                     convergence_tol=solver["convergence_tol"],
                                     divergence_tol=solver["divergence_tol"],
                                     divergence_threshold=divergence_threshold, seed=seed, counter=counter)
            data = logdensity_fn.data(dtype=torch.float64, num_chains=C)
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

        return SamplingAlgorithm(init_fn, step_fn)


__all__ = ["nuts", "build_kernel", "NUTSInfo", "NUTSNoise", "init"]
