"""Dense linear algebra with JAX's failure semantics (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``jax.scipy.linalg.cholesky`` / ``jnp.linalg.solve`` / ``lu_factor`` return
NaN-filled results on a non-positive-definite or singular input instead of
raising; torch raises.  These wrappers use torch's ``*_ex`` forms and map a
failed factorisation to NaN so that a divergent trajectory flows into a NaN
energy and a rejection, as it does in the reference
(geomjax/mcmc/proposal.py:107-108).
"""
import torch


def cholesky(A):
    L, info = torch.linalg.cholesky_ex(A)
    return torch.where(info == 0, L, torch.full_like(L, float("nan")))


def solve(A, b):
    x, info = torch.linalg.solve_ex(A, b)
    return torch.where(info == 0, x, torch.full_like(x, float("nan")))


def lu_factor(A):
    LU, piv, info = torch.linalg.lu_factor_ex(A)
    return torch.where(info == 0, LU, torch.full_like(LU, float("nan"))), piv
