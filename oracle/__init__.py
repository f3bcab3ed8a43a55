"""CPU float64 restatement of geomjax's Riemannian integrator hot path.

TEST INFRASTRUCTURE ONLY. This package is the correctness oracle for the HIP
kernels in ``geomjax_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only to compare results with
(or, for ``bench.py``, as the timed CPU baseline).  The product path never
imports, calls or falls back to anything in here.

What it is
----------
A line-by-line restatement, in PyTorch on the CPU in float64, of the reference
algorithm as the reference writes it: dense metric matrices, dense Cholesky /
LU / slogdet, and automatic differentiation (``torch.func.grad`` / ``jvp`` /
``jacfwd`` / ``vmap`` standing where the reference uses ``jax.grad`` /
``jax.jvp`` / ``jax.jacfwd`` / ``jax.vmap``).  The HIP kernels use closed-form
O(D) Sherman-Morrison algebra and weighted-Gram contractions instead, so the two
are independent computations of the same numbers.

The reference's random keys are replaced by host-fed noise: the standard-normal
vector ``z`` that ``generate_gaussian_noise`` would draw
(geomjax/util.py:58-83) and the uniform ``u`` behind ``jax.random.bernoulli``
(geomjax/mcmc/proposal.py:177-178; bernoulli(key, p) is ``uniform(key) < p``).

Modules
-------
models        target log densities and metric functions (funnel, logistic)
mcmc          proposal / accept tail, static integration, Euclidean HMC
rmhmc         implicit midpoint RMHMC (geomjax/rmhmc/*)
lmcmonge      Lagrangian Monte Carlo with the Monge metric (geomjax/lmcmonge/*)
lmcmc         Lagrangian Monte Carlo with a general metric (geomjax/lmcmc/*)
mmala         manifold MALA with a dense metric (geomjax/lmcmc/mmala.py, diffusions.py)
diagnostics   ESS and R-hat in numpy (geomjax/diagnostics.py)

How each function is pinned (checks that do not just retype its formula) is
listed in tests/test_oracle_*.py.  Functions with no such check are marked
"parity unpinned" in their docstring.

Everything here is float64 and CPU; tensors are created with explicit dtype so
importing the oracle changes no global torch setting.
"""

import torch

DTYPE = torch.float64
