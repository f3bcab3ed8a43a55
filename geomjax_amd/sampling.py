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
