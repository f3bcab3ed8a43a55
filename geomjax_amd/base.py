"""Algorithm containers shared by every sampler (geomjax/base.py:88-113)."""
from typing import Callable, NamedTuple


class SamplingAlgorithm(NamedTuple):
    """A pair of functions ``init(position)`` and ``step(rng_key, state)``.

    Same container as the reference's ``SamplingAlgorithm``
    (geomjax/base.py:88-113).  States and infos here hold batched device
    tensors: the leading axis is the chain axis that the reference gets from
    ``jax.vmap``.
    """

    init: Callable
    step: Callable
