"""Gradients for tuning the step size (and alpha2 of the Monge metric) of RMHMC by gradient.

The quantity tuned is the expected squared jump distance times acceptance,
mean over chains of  p_c * |q_L,c - q_0,c|^2  with p_c = min(1, e^dH_c), the
criterion of the reference's ChEES-style adaptation
(geomjax/adaptation/chees_adaptation_lmcmonge.py:222-257 uses the proposal
derivative with respect to alpha2 from jax.jacfwd, :475-519).

Derivatives come from the HIP tangent kernel: one transition evaluated on
dual numbers seeded with d/d log eps and d/d log alpha2, i.e. the exact
derivative of the executed computation, fixed-point iterates included (the
semantics of jax.jacfwd through the reference integrator).  With two
parameters forward tangents cost about two primal transitions and keep no
trajectory in memory, so no checkpointing is needed; the momentum is
reparameterised as chol(G(q0; alpha2)) z with z held fixed.

Two kernels: RMHMC-Monge on Neal's funnel (derivatives in log step size and
log alpha2) and RMHMC-Fisher on logistic regression (log step size; the
Fisher metric has no alpha2, its derivative column is zero).

``ESJDAcceptance`` exposes it as a ``torch.autograd.Function`` of
(log_step_size, log_alpha2); ``adapt`` runs a small Adam loop on them.
"""
import math
from typing import NamedTuple

import torch

from . import _lib
from .metrics import FisherMetric, MongeMetric
from .models import LogisticBatch, LogisticRegression, NealFunnel, logistic_params
from .random import HostNoise, PRNGKey
from .rmhmc import _params, _solver_kwargs, implicit_midpoint


class TangentResult(NamedTuple):
    objective: torch.Tensor      # [C] acceptance * squared jump
    acceptance: torch.Tensor     # [C]
    sq_jump: torch.Tensor        # [C]
    d_objective: torch.Tensor    # [C, 2]: d/d log step_size, d/d log alpha2
    d_acceptance: torch.Tensor   # [C, 2]
    d_sq_jump: torch.Tensor      # [C, 2]
    num_fevals: torch.Tensor     # [C]


def esjd_acceptance(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, *,
                    integrator=implicit_midpoint, lanes_per_chain=0) -> TangentResult:
    """Per-chain objective and its derivatives w.r.t. (log step_size, log alpha2).

    Funnel + MongeMetric: both columns; a float64 state runs the check path, a
    float32 state the fast path (float32 trajectory and tangents, float64
    energies, as the tuned sampler).  LogisticRegression / LogisticBatch +
    FisherMetric: column 0 (column 1 is zero); a float64 state runs the check
    path (D <= 64), a float32 state the fast path (D <= 64: the tuned kernel's
    float32 state and split-bf16 MFMA passes with tangents; 65..128: a float32
    form of the check path on fp32 MFMA; float64 energies in both)."""
    funnel = isinstance(logdensity_fn, NealFunnel) and isinstance(metric_fn, MongeMetric)
    logistic = isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) and isinstance(metric_fn, FisherMetric)
    if not (funnel or logistic):
        raise NotImplementedError("tangent kernels: RMHMC-Monge on the funnel, RMHMC-Fisher on logistic regression")
    solver = _solver_kwargs(integrator)
    q = state.position
    if q.dim() != 2 or q.dtype not in (torch.float64, torch.float32) or not q.is_cuda:
        raise ValueError("position must be a [C, D] float64 or float32 tensor on a HIP device")
    q = q.contiguous()
    C, D = q.shape
    dev = q.device
    logp = state.logdensity.reshape(C).to(torch.float64).contiguous()
    if isinstance(rng_key, HostNoise):
        z = rng_key.z.reshape(C, D).to(device=dev, dtype=q.dtype).contiguous()
        seed, counter = 0, 0
    elif isinstance(rng_key, PRNGKey):
        z, seed, counter = None, rng_key.seed, rng_key.counter
    else:
        raise TypeError("rng_key must be HostNoise or PRNGKey")
    f = lambda *s: torch.empty(*s, dtype=torch.float64, device=dev)
    out = TangentResult(f(C), f(C), f(C), f(C, 2), f(C, 2), f(C, 2), torch.empty(C, dtype=torch.int64, device=dev))
    io = _lib.TangentIO(position=q.data_ptr(), logdensity=logp.data_ptr(), z=_lib.ptr(z),
                        objective=out.objective.data_ptr(), acceptance=out.acceptance.data_ptr(),
                        sq_jump=out.sq_jump.data_ptr(), d_objective=out.d_objective.data_ptr(),
                        d_acceptance=out.d_acceptance.data_ptr(), d_sq_jump=out.d_sq_jump.data_ptr(),
                        num_fevals=out.num_fevals.data_ptr())
    if funnel:
        code = _lib.GJX_F64 if q.dtype == torch.float64 else _lib.GJX_F32
        params = _params(C, D, code, step_size, num_integration_steps, solver, metric_fn.alpha2,
                         logdensity_fn.sigma, 1000.0, lanes_per_chain=lanes_per_chain, seed=seed, counter=counter)
        rc = _lib.load().gjx_rmhmc_monge_funnel_tangent(params, io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_rmhmc_monge_funnel_tangent")
    else:
        if D != logdensity_fn.dim:
            raise ValueError(f"positions have {D} coordinates, the model {logdensity_fn.dim}")
        params = logistic_params(logdensity_fn, C, L=num_integration_steps, step_size=step_size,
                                 max_iters=solver["max_iters"], convergence_tol=solver["convergence_tol"],
                                 divergence_tol=solver["divergence_tol"], seed=seed, counter=counter,
                                 dtype=q.dtype)
        data = logdensity_fn.data(dtype=q.dtype, num_chains=C)
        rc = _lib.load().gjx_rmhmc_fisher_logistic_tangent(params, data, io, _lib.stream_handle(dev))
        _lib.check(rc, "gjx_rmhmc_fisher_logistic_tangent")
    return out


def _metric_for(logdensity_fn, alpha2):
    return FisherMetric() if isinstance(logdensity_fn, (LogisticRegression, LogisticBatch)) else MongeMetric(alpha2)


class ESJDAcceptance(torch.autograd.Function):
    """mean_c p_c |q_L,c - q_0,c|^2 as a differentiable function of (log eps, log alpha2).

    ``apply(log_step_size, log_alpha2, state, rng_key, logdensity_fn,
    num_integration_steps, integrator)``; gradients flow to the first two
    (0-dim tensors).
    """

    @staticmethod
    def forward(ctx, log_step_size, log_alpha2, state, rng_key, logdensity_fn, num_integration_steps,
                integrator=implicit_midpoint):
        eps = math.exp(float(log_step_size))
        a2 = math.exp(float(log_alpha2))
        r = esjd_acceptance(rng_key, state, logdensity_fn, eps, _metric_for(logdensity_fn, a2),
                            num_integration_steps, integrator=integrator)
        grad = r.d_objective.mean(dim=0)
        ctx.save_for_backward(grad.to(log_step_size.device))
        return r.objective.mean().to(log_step_size.device)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad[0], g * grad[1], None, None, None, None, None


def adapt(rng_key, state, logdensity_fn, num_integration_steps, *, step_size=0.1, alpha2=1.0, num_steps=50,
          learning_rate=0.05, integrator=implicit_midpoint):
    """Gradient ascent (Adam) on (log eps, log alpha2) of the ESJD x acceptance criterion.

    Each iteration draws fresh momenta (device Philox, key counter advanced),
    evaluates the criterion and its tangent derivative, takes an Adam step and
    moves the chains by one transition with the current parameters.  Returns
    (step_size, alpha2, state, history).
    """
    from .rmhmc import rmhmc

    theta = torch.tensor([math.log(step_size), math.log(alpha2)], dtype=torch.float64)
    m = torch.zeros(2, dtype=torch.float64)
    v = torch.zeros(2, dtype=torch.float64)
    b1, b2, tiny = 0.9, 0.999, 1e-8
    hist = []
    key = rng_key
    for it in range(1, num_steps + 1):
        eps, a2 = math.exp(theta[0]), math.exp(theta[1])
        r = esjd_acceptance(PRNGKey(key.seed, key.counter + it), state, logdensity_fn, eps,
                            _metric_for(logdensity_fn, a2), num_integration_steps, integrator=integrator)
        g = r.d_objective.mean(dim=0).cpu()
        if not torch.all(torch.isfinite(g)):
            g = torch.zeros(2, dtype=torch.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        theta = theta + learning_rate * (m / (1 - b1 ** it)) / (torch.sqrt(v / (1 - b2 ** it)) + tiny)
        hist.append((eps, a2, float(r.objective.mean()), float(r.acceptance.mean())))
        alg = rmhmc(logdensity_fn, eps, _metric_for(logdensity_fn, a2), num_integration_steps, integrator=integrator)
        state, _ = alg.step(PRNGKey(key.seed ^ 0x5DEECE66D, key.counter + it), state)
    return math.exp(theta[0]), math.exp(theta[1]), state, hist
