"""NUTS with the LMC-Monge integrator (geomjax/lmcmonge/nuts.py), on the HIP kernel.

Same API as the reference: ``nutslmc(logdensity_fn, step_size,
inverse_mass_matrix, *, alpha2=0.001, max_num_doublings=10,
divergence_threshold=1000, integrator=lan_integrator)`` with ``init`` (=
lmcmonge.init, the ``LMCState``) and ``step``; the kernel
``build_kernel(integrator, divergence_threshold, max_num_doublings)(rng_key,
state, logdensity_fn, step_size, inverse_mass_matrix, alpha2,
stopping_criterion="euc")`` returns ``(LMCState, NUTSInfo)``.  The whole
transition (velocity draw, iterative doubling over lan_integrator steps,
progressive multinomial sampling with the energy -l + K - J, U-turn checks)
runs in one HIP kernel, one wave per chain, on Neal's funnel: float64 for a
float64 state, the tuned form for a float32 one (float32 trajectory, log
density, kinetic-energy sums, volume change and tree weights in float64).

``rng_key`` is a ``PRNGKey`` (device Philox) or ``geomjax_amd.nuts.NUTSNoise``
(the velocity normals, direction / merge uniforms per expansion, progressive
uniforms per integration step).  ``reference_faithful`` selects the
reference's algebra or the exact one, as for ``lmcmonge.lmc``.
``NUTSInfo.trajectory_leftmost_state`` / ``trajectory_rightmost_state`` carry
the positions of the trajectory's two ends.
"""
from typing import NamedTuple

import torch

from . import _lib
from .lmcmonge import LMCState, _check, _params, init, lan_integrator
from .metrics import _check_alpha2
from .nuts import NUTSNoise
from .random import PRNGKey

__all__ = ["NUTSInfo", "build_kernel", "nutslmc", "init"]


class NUTSInfo(NamedTuple):  # geomjax/lmcmonge/nuts.py:39-77 (same fields and order)
    velocity: torch.Tensor
    is_divergent: torch.Tensor
    is_turning: torch.Tensor
    energy: torch.Tensor
    trajectory_leftmost_state: torch.Tensor
    trajectory_rightmost_state: torch.Tensor
    num_trajectory_expansions: torch.Tensor
    num_integration_steps: torch.Tensor
    acceptance_rate: torch.Tensor


def build_kernel(integrator=lan_integrator, divergence_threshold: int = 1000, max_num_doublings: int = 10, *,
                 reference_faithful=True):
    """geomjax/lmcmonge/nuts.py:80-207."""
    if integrator is not lan_integrator:
        raise NotImplementedError("only lan_integrator is implemented on the GPU")
    M = int(max_num_doublings)

    def kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, alpha2, stopping_criterion="euc"):
        _check_alpha2(alpha2)
        if stopping_criterion not in _lib.GJX_UTURN:
            raise ValueError(f"stopping_criterion must be one of {sorted(_lib.GJX_UTURN)} (got {stopping_criterion!r})")
        single = state.position.dim() == 1
        position = (state.position[None] if single else state.position).contiguous()
        if position.dtype not in (torch.float64, torch.float32):
            raise TypeError("NUTS-LMC positions must be float64 (plain) or float32 (tuned)")
        im = _check(logdensity_fn, position, inverse_mass_matrix)
        dev = position.device
        C, D = position.shape
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
        io = _lib.LmcNutsIO(inverse_mass_matrix=im.data_ptr(), position=position.data_ptr(),
                            logdensity=logdensity.data_ptr(), volume_adjustment=J.data_ptr(), z=_lib.ptr(noise[0]),
                            u_direction=_lib.ptr(noise[1]), u_merge=_lib.ptr(noise[2]), u_step=_lib.ptr(noise[3]),
                            **{k: v.data_ptr() for k, v in out.items()})
        code = _lib.GJX_F64 if position.dtype == torch.float64 else _lib.GJX_F32
        params = _params(C, D, code, step_size, 0, alpha2, logdensity_fn.sigma, divergence_threshold,
                         reference_faithful, seed=seed, counter=counter)
        rc = _lib.load().gjx_lmcmonge_funnel_nuts(params, M, _lib.GJX_UTURN[stopping_criterion], io,
                                                  _lib.stream_handle(dev))
        _lib.check(rc, "gjx_lmcmonge_funnel_nuts")
        sq = (lambda t: t[0]) if single else (lambda t: t)
        new_state = LMCState(sq(out["new_position"]), sq(out["new_logdensity"]), sq(out["new_logdensity_grad"]),
                             sq(out["new_volume_adjustment"]))
        info = NUTSInfo(sq(out["velocity"]), sq(out["is_divergent"].bool()), sq(out["is_turning"].bool()),
                        sq(out["energy"]), sq(out["leftmost_position"]), sq(out["rightmost_position"]),
                        sq(out["num_trajectory_expansions"]), sq(out["num_integration_steps"]),
                        sq(out["acceptance_rate"]))
        return new_state, info

    return kernel


class nutslmc:
    """geomjax/lmcmonge/nuts.py:210-297 (exported as ``nutslmcmonge``)."""

    init = staticmethod(init)
    build_kernel = staticmethod(build_kernel)

    def __new__(cls, logdensity_fn, step_size, inverse_mass_matrix, *, alpha2: float = 0.001,
                max_num_doublings: int = 10, divergence_threshold: int = 1000, integrator=lan_integrator,
                reference_faithful=True):
        from .base import SamplingAlgorithm

        kernel = cls.build_kernel(integrator, divergence_threshold, max_num_doublings,
                                  reference_faithful=reference_faithful)

        def init_fn(position):
            return cls.init(position, logdensity_fn)

        def step_fn(rng_key, state):
            return kernel(rng_key, state, logdensity_fn, step_size, inverse_mass_matrix, alpha2)

        return SamplingAlgorithm(init_fn, step_fn)
