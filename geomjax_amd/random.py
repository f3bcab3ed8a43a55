"""Randomness for the samplers.

The reference draws from JAX keys (``jax.random.split/normal/bernoulli``,
geomjax/rmhmc/rmhmc.py:158, geomjax/util.py:82, geomjax/mcmc/proposal.py:178).
Here a kernel call takes one of

* ``HostNoise(z, u)``: the standard-normal vector ``z`` [C, D] that
  ``generate_gaussian_noise`` would draw for the momentum (velocity) and the
  uniform ``u`` [C] behind ``jax.random.bernoulli(key, p)`` (= ``u < p``),
  fed from the caller.  This is the parity mode: given the same z and u the
  result equals the reference's.
* ``PRNGKey``: a (seed, counter) pair for the in-kernel Philox4x32-10
  generator.  Every chain and transition gets its own counter block, so one
  key per call covers all chains.  Not threefry-identical to JAX.
"""
from typing import NamedTuple

import torch


class HostNoise(NamedTuple):
    z: torch.Tensor
    u: torch.Tensor


class PRNGKey(NamedTuple):
    seed: int
    counter: int = 0


_MASK64 = (1 << 64) - 1


def key(seed: int) -> PRNGKey:
    return PRNGKey(int(seed) & _MASK64, 0)


def _mix(x: int) -> int:
    # splitmix64 finaliser: decorrelates derived seeds
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _MASK64
    return x ^ (x >> 31)


def split(k: PRNGKey, num: int = 2):
    """``num`` independent keys derived from ``k`` (as jax.random.split)."""
    return [PRNGKey(_mix(k.seed ^ _mix(k.counter * 1000003 + i + 1)), 0) for i in range(num)]


def fold_in(k: PRNGKey, transitions: int) -> PRNGKey:
    """The key after ``transitions`` transitions drawn from ``k``."""
    return PRNGKey(k.seed, k.counter + int(transitions))


def host_noise(num_chains, dim, *, generator=None, dtype=torch.float64, device="cpu"):
    """Draw one transition's HostNoise on the host (numpy-free helper)."""
    z = torch.randn(num_chains, dim, generator=generator, dtype=dtype)
    u = torch.rand(num_chains, generator=generator, dtype=torch.float64)
    return HostNoise(z.to(device), u.to(device))
