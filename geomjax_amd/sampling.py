"""Multi-chain, multi-transition driver.

``run_chains`` is the reference's user loop ``jax.lax.scan`` over
``jax.vmap(sampler.step)`` (examples/funnel/main.py:7-25) as one persistent
HIP launch: every chain's state stays in registers across all transitions and
only the samples (and per-transition acceptance probabilities) are written.
"""
from typing import Any, NamedTuple, Optional

import torch

from .base import SamplingAlgorithm


class SampleResult(NamedTuple):
    state: Any                       # final chain state (the sampler's State tuple)
    samples: Optional[torch.Tensor]  # [T, C, D] positions after each transition
    acceptance_rate: torch.Tensor    # [T, C] float32
    num_accepted: torch.Tensor       # [C] int32
    num_fevals: torch.Tensor         # [C] int64: vector-field evaluations executed


class BoundAlgorithm(SamplingAlgorithm):
    """SamplingAlgorithm that also knows its persistent multi-transition launch."""

    def __new__(cls, init, step, sample_fn):
        self = super().__new__(cls, init, step)
        self.sample_fn = sample_fn
        return self


def run_chains(algorithm, initial_state, num_samples, rng_key, **kw) -> SampleResult:
    """Run ``num_samples`` transitions of every chain in ``initial_state``."""
    fn = getattr(algorithm, "sample_fn", None)
    if fn is None:
        raise TypeError("algorithm has no persistent sample launch (build it with geomjax_amd.rmhmc/lmcmonge/...)")
    return fn(rng_key, initial_state, num_samples, **kw)


class GraphSampler:
    """``num_samples`` transitions of a bound sampler captured once into a HIP graph.

    The reference re-dispatches its jitted scan for every chunk of samples; here
    one capture holds the chunk's whole device work -- the draw of its standard
    normals and uniforms into fixed buffers (torch's generator, whose offsets
    advance on every replay), the persistent sample launch on them (``HostNoise``)
    and the copy of the final state back into fixed state buffers -- and each
    ``replay()`` continues the chains by one chunk with a single graph launch.

    With ``draw_noise=False`` the caller fills ``z`` [T, C, D] and ``u`` [T, C]
    before each replay (the host-fed parity mode).  Lazy set-up (the packed
    design matrix and the workspace of the float32 logistic path, library load)
    happens in an eager run before the capture; ``state`` is copied into the
    sampler's own buffers (``self.state``), which every replay overwrites.
    """

    def __init__(self, algorithm, state, num_samples, *, draw_noise=True, store_samples=True):
        from .random import HostNoise

        fn = getattr(algorithm, "sample_fn", None)
        if fn is None:
            raise TypeError("algorithm has no persistent sample launch")
        pos = state.position
        if pos.dim() != 2 or not pos.is_cuda:
            raise ValueError("position must be a [C, D] tensor on a HIP device")
        T = int(num_samples)
        C, D = pos.shape
        dev = pos.device
        self.num_samples = T
        self.z = torch.zeros(T, C, D, dtype=pos.dtype, device=dev)
        self.u = torch.zeros(T, C, dtype=torch.float64, device=dev)
        self.state = type(state)(*(torch.as_tensor(t, device=dev).clone() for t in state))
        self.samples = torch.empty(T, C, D, dtype=pos.dtype, device=dev) if store_samples else None
        noise = HostNoise(self.z, self.u)
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            fn(noise, self.state, T, samples_out=self.samples, store_samples=store_samples)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            if draw_noise:
                self.z.normal_()
                self.u.uniform_()
            r = fn(noise, self.state, T, samples_out=self.samples, store_samples=store_samples)
            for dst, src in zip(self.state, r.state):
                dst.copy_(src)
        self.acceptance_rate = r.acceptance_rate  # [T, C], rewritten by every replay
        self.num_accepted = r.num_accepted
        self.num_fevals = r.num_fevals

    def replay(self):
        """Advance every chain by ``num_samples`` transitions; returns the state buffers."""
        self.graph.replay()
        return self.state
