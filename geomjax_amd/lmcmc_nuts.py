"""NUTS with the general-metric LMC integrator (geomjax/lmcmc/nuts.py), on the HIP kernel.

Same API as the reference: ``nuts(logdensity_fn, step_size, metric_fn, *,
max_num_doublings=10, divergence_threshold=1000, integrator=lan_integrator)``
(exported as ``nutslmc``) with ``init`` (= lmcmc.init, the ``LMCState``) and
``step``; the kernel ``build_kernel(integrator, divergence_threshold)(rng_key,
state, logdensity_fn, step_size, metric_fn, max_num_doublings=10,
stopping_criterion="euc")`` returns ``(LMCState, NUTSInfo)``.  The whole
transition (velocity draw v0 = L^-T z, iterative doubling over lan_integrator
steps, progressive multinomial sampling with the energy -l - log det G / 2 +
v^T G v / 2 - J, U-turn checks) runs in one HIP kernel: Bayesian logistic
regression with the Fisher metric, one posterior or a ``LogisticBatch``, every
posterior with D <= 64; float64 for a float64 state (one 256-lane workgroup per
chain), the tuned form for a float32 one (one wave per chain, split-bf16 MFMA
Gram passes, float64 log densities, log-determinants, volume change, energies
and tree weights).

``rng_key`` is a ``PRNGKey`` (device Philox) or ``geomjax_amd.nuts.NUTSNoise``.
``NUTSInfo.trajectory_leftmost_state`` / ``trajectory_rightmost_state`` carry
the positions of the trajectory's two ends.
"""
from typing import NamedTuple

import torch

from . import _lib
from .lmcmc import LMCState, _check, init, lan_integrator
from .models import logistic_params
from .nuts import NUTSNoise
from .random import PRNGKey

__all__ = ["NUTSInfo", "build_kernel", "nuts", "init"]


class NUTSInfo(NamedTuple):  # geomjax/lmcmc/nuts.py:36-77 (same fields and order)
    velocity: torch.Tensor
    is_divergent: torch.Tensor
    is_turning: torch.Tensor
    energy: torch.Tensor
    trajectory_leftmost_state: torch.Tensor
    trajectory_rightmost_state: torch.Tensor
    num_trajectory_expansions: torch.Tensor
    num_integration_steps: torch.Tensor
    acceptance_rate: torch.Tensor


def build_kernel(integrator=lan_integrator, divergence_threshold: int = 1000):
    """geomjax/lmcmc/nuts.py:78-161."""
    if integrator is not lan_integrator:
        raise NotImplementedError("only lan_integrator is implemented on the GPU")

    def kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings: int = 10,
               stopping_criterion: str = "euc"):
        if stopping_criterion not in _lib.GJX_UTURN:
            raise ValueError(f"stopping_criterion must be one of {sorted(_lib.GJX_UTURN)} (got {stopping_criterion!r})")
        from .lmcmc_dense import is_dense_metric_fn

        if is_dense_metric_fn(logdensity_fn, metric_fn):
            return _dense(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings, stopping_criterion,
                          divergence_threshold)
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        if position.dtype not in (torch.float64, torch.float32):
            raise TypeError("NUTS-LMC positions must be float64 (plain) or float32 (tuned)")
        _check(logdensity_fn, metric_fn, position)
        dev = position.device
        C, D = position.shape
        M = int(max_num_doublings)
        logdensity = state.logdensity.reshape(C).to(torch.float64).contiguous()
        J = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).contiguous()
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
        out = dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), new_volume_adjustment=sca(),
                   velocity=vec(), energy=sca(), acceptance_rate=sca(), is_divergent=sca(torch.uint8),
                   is_turning=sca(torch.uint8), num_trajectory_expansions=sca(torch.int32),
                   num_integration_steps=sca(torch.int32), leftmost_position=vec(), rightmost_position=vec())
        io = _lib.LmcNutsIO(position=position.data_ptr(), logdensity=logdensity.data_ptr(),
                            volume_adjustment=J.data_ptr(), z=_lib.ptr(noise[0]), u_direction=_lib.ptr(noise[1]),
                            u_merge=_lib.ptr(noise[2]), u_step=_lib.ptr(noise[3]),
                            **{k: v.data_ptr() for k, v in out.items()})
        params = logistic_params(logdensity_fn, C, step_size=step_size, divergence_threshold=divergence_threshold,
                                 seed=seed, counter=counter, dtype=position.dtype)
        data = logdensity_fn.data(dtype=position.dtype, num_chains=C)
        rc = _lib.load().gjx_lmc_fisher_logistic_nuts(params, data, M, _lib.GJX_UTURN[stopping_criterion], io,
                                                      _lib.stream_handle(dev))
        _lib.check(rc, "gjx_lmc_fisher_logistic_nuts")
        sq = (lambda t: t[0]) if single else (lambda t: t)
        new_state = LMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]),
                             sq(out["new_volume_adjustment"]))
        info = NUTSInfo(sq(out["velocity"]), sq(out["is_divergent"].bool()), sq(out["is_turning"].bool()),
                        sq(out["energy"]), sq(out["leftmost_position"]), sq(out["rightmost_position"]),
                        sq(out["num_trajectory_expansions"]), sq(out["num_integration_steps"]),
                        sq(out["acceptance_rate"]))
        return new_state, info

    return kernel


def _dense(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings, stopping_criterion,
           divergence_threshold):
    """A user dense metric: the batched expansion of ``nuts_dense`` over the
    dense-metric LMC steps."""
    from . import nuts_dense

    single = state.position.dim() == 1
    st = LMCState(*((torch.as_tensor(t)[None] if single else t) for t in state))
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
    q, l, g, J, inf = nuts_dense.lmc_transition(noise, st, logdensity_fn, step_size, metric_fn, M,
                                                stopping_criterion, divergence_threshold)
    sq = (lambda t: t[0]) if single else (lambda t: t)
    info = NUTSInfo(sq(inf["momentum"]), sq(inf["is_divergent"]), sq(inf["is_turning"]), sq(inf["energy"]),
                    sq(inf["left"]), sq(inf["right"]), sq(inf["num_trajectory_expansions"]),
                    sq(inf["num_integration_steps"]), sq(inf["acceptance_rate"]))
    return LMCState(sq(q), sq(l), sq(g), sq(J)), info


class nuts:
    """geomjax/lmcmc/nuts.py:164-250 (exported as ``nutslmc``)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, metric_fn, max_num_doublings: int = 10,
                divergence_threshold: int = 1000, integrator=lan_integrator):
        from .base import SamplingAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, metric_fn, max_num_doublings)

        return SamplingAlgorithm(init_fn, step_fn)
