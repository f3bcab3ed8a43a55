"""Target densities served by the HIP kernels.

The reference takes arbitrary JAX callables ``logdensity_fn`` and
``metric_fn`` and differentiates them with ``jax.grad``/``jvp``/``jacfwd``.
A GPU kernel cannot trace a Python callable, so the boundary accepts a closed
set of model descriptors whose log density, gradient and Hessian-vector
products are written in closed form inside the kernels.
"""
import torch

from . import _lib


class NealFunnel:
    """Neal's funnel as in examples/funnel/main.py:28-54.

    The log-scale coordinate is the LAST one: v = theta[D-1] ~ N(0, sigma^2),
    x_i | v ~ N(0, e^v) for i < D-1.  ``mean`` is accepted for signature
    parity and must be 0 (the reference stores it but never uses it).
    """

    kind = "funnel"

    def __init__(self, D=2, mean=0.0, sigma=3.0):
        if D < 2:
            raise ValueError(f"the funnel needs D >= 2 (got {D})")
        if mean != 0.0:
            raise ValueError("NealFunnel.mean is unused by the reference's logp and must be 0")
        if not sigma > 0:
            raise ValueError(f"sigma must be > 0 (got {sigma})")
        self.D = int(D)
        self.mean = float(mean)
        self.sigma = float(sigma)

    def value_and_grad(self, position):
        """(logdensity [C], grad [C, D]) on the device of ``position`` via HIP."""
        return funnel_value_and_grad(position, self.sigma)

    def __repr__(self):
        return f"NealFunnel(D={self.D}, sigma={self.sigma})"


def _dtype_code(t):
    if t.dtype == torch.float64:
        return _lib.GJX_F64
    if t.dtype == torch.float32:
        return _lib.GJX_F32
    raise TypeError(f"positions must be float64 or float32 (got {t.dtype})")


def funnel_value_and_grad(position, sigma):
    if not position.is_cuda:
        raise ValueError("positions must be on a HIP device (cuda:N)")
    pos = position.contiguous()
    squeeze = pos.dim() == 1
    if squeeze:
        pos = pos[None, :]
    C, D = pos.shape
    logp = torch.empty(C, dtype=torch.float64, device=pos.device)
    grad = torch.empty_like(pos)
    lib = _lib.load()
    rc = lib.gjx_funnel_value_and_grad(C, D, float(sigma), _dtype_code(pos), pos.data_ptr(), logp.data_ptr(),
                                       grad.data_ptr(), _lib.stream_handle(pos.device))
    _lib.check(rc, "gjx_funnel_value_and_grad")
    if squeeze:
        return logp[0], grad[0]
    return logp, grad
