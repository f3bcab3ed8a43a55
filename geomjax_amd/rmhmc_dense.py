"""RMHMC with a general dense user metric (geomjax/rmhmc with any ``metric_fn``).

The user's ``logdensity_fn`` and ``metric_fn`` (torch operations that
``torch.func`` can transform, on the device of their input) are evaluated and
differentiated by ``torch.func`` on the GPU, batched over chains; the vector
field of the implicit midpoint -- the Cholesky factor of the metric, u = G^-1 p,
dT/dq = tr(G^-1 dG)/2 - u^T dG u/2, log det G, the momentum draw chol(G) z --
runs in ``gjx_dense_rmhmc_field``.  The fixed-point iteration follows the
reference's per-chain stop rule (geomjax/rmhmc/integrators.py:53-89): chains
freeze independently, and the loop ends when no chain is active (eagerly) or
after ``max_iters`` evaluations (under HIP-graph capture, where reading the
flags back is not possible; frozen chains do not change, so both give the same
result).  float64, D <= 64.  Draws as in ``lmcmc_dense``.
"""
import math

import torch

from . import _lib
from .lmcmc_dense import _check, _metric, _noise, _value_and_grad
from .random import HostNoise, PRNGKey

__all__ = ["init", "transition", "sample"]


def _field(G, dG=None, p=None, z=None):
    """(p, u, dT/dq, log det G, p^T u) of the dense-metric kernel."""
    C, D = G.shape[0], G.shape[1]
    dev = G.device
    vec = lambda: torch.empty(C, D, dtype=torch.float64, device=dev)
    u, pout = vec(), vec()
    dtdq = vec() if dG is not None else None
    logdet = torch.empty(C, dtype=torch.float64, device=dev)
    pu = torch.empty(C, dtype=torch.float64, device=dev)
    io = _lib.DenseMetricIO(num_chains=C, dim=D, step_size=0.0, metric=_lib.ptr(G), metric_jacobian=_lib.ptr(dG),
                            momentum=_lib.ptr(p), z=_lib.ptr(z), velocity_out=u.data_ptr(),
                            momentum_out=pout.data_ptr(), dtdq=_lib.ptr(dtdq), logdet=logdet.data_ptr(),
                            pu=pu.data_ptr())
    _lib.check(_lib.load().gjx_dense_rmhmc_field(io, _lib.stream_handle(dev)), "gjx_dense_rmhmc_field")
    return pout, u, dtdq, logdet, pu


def init(position, logdensity_fn):
    """geomjax/rmhmc/rmhmc.py:96-98 for a user log density."""
    from .rmhmc import RMHMCState

    single = position.dim() == 1
    q = (position[None] if single else position).contiguous()
    _check(q)
    _lib.load()  # the sampler's kernels are native: fail here, not at the first step, without them
    l, g = _value_and_grad(logdensity_fn, q)
    if single:
        return RMHMCState(position, l[0], g[0])
    return RMHMCState(position, l, g)


def _kinetic(logdet, pu, D):
    """-log N(p; 0, G) = p^T G^-1 p / 2 + log det G / 2 + (D/2) log 2 pi (metrics.py:66-69)."""
    return 0.5 * pu + 0.5 * logdet + 0.5 * D * math.log(2.0 * math.pi)


def _implicit_midpoint_step(q0, p0, eps, logdensity_fn, metric_fn, solver):
    """geomjax/rmhmc/integrators.py:92-156 for every chain: the fixed point of
    x = z0 + (eps/2) F(x), then one explicit half step from the last iterate.
    ``eps`` is a float or a [C] tensor of per-chain signed step sizes."""
    h = (0.5 * eps)[:, None] if torch.is_tensor(eps) else 0.5 * eps

    def F(q, p, initial):
        _, g = _value_and_grad(logdensity_fn, q)
        G, dG = _metric(metric_fn, q)
        _, u, dtdq, _, _ = _field(G, dG, p=p)
        return initial[0] + h * u, initial[1] - h * (dtdq - g)

    def max_abs(a, b):
        return torch.maximum((a[0] - b[0]).abs().amax(dim=1), (a[1] - b[1]).abs().amax(dim=1))

    x0 = (q0, p0)
    x = F(q0, p0, x0)
    norm = max_abs(x, x0)
    n = torch.zeros_like(norm, dtype=torch.int64)
    max_iters = int(solver["max_iters"])
    tol, div = float(solver["convergence_tol"]), float(solver["divergence_tol"])
    capturing = torch.cuda.is_current_stream_capturing()
    for _ in range(max_iters):
        active = (n < max_iters) & torch.isfinite(norm) & (norm < div) & (norm > tol)
        if not capturing and not bool(active.any()):
            break
        xn = F(x[0], x[1], x0)
        normn = max_abs(xn, x)
        m = active[:, None]
        x = (torch.where(m, xn[0], x[0]), torch.where(m, xn[1], x[1]))
        norm = torch.where(active, normn, norm)
        n = n + active.to(torch.int64)
    success = torch.isfinite(norm) & (norm <= tol)
    q1, p1 = F(x[0], x[1], x)
    return q1, p1, (success, norm, n)


def _per_chain_steps(num_integration_steps, C, dev, max_num_integration_steps):
    """(L, per-chain counts or None): an int, or a [C] integer tensor (the dynamic
    kernels' form) clamped to ``max_num_integration_steps`` when given."""
    if isinstance(num_integration_steps, torch.Tensor):
        Ls = num_integration_steps.to(device=dev, dtype=torch.int64).reshape(C)
        L = int(max_num_integration_steps) if max_num_integration_steps is not None else int(Ls.max().item())
        L = max(L, 0)
        return L, Ls.clamp(0, L)
    return int(num_integration_steps), None


def transition(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, solver,
               divergence_threshold=1000.0, max_num_integration_steps=None):
    """One RMHMC transition (geomjax/rmhmc/rmhmc.py:131-174, 376-462) of every chain;
    ``num_integration_steps`` an int or a [C] tensor of per-chain counts (a chain
    whose count is reached keeps its state while the others step on).
    Returns (RMHMCState, RMHMCInfo, per-step FixedPointIterationInfo list)."""
    from .proposal import Proposal
    from .rmhmc import IntegratorState, RMHMCInfo, RMHMCState
    from .lmcmc_dense import resolve_metric_fn

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)

    q0 = state.position.contiguous()
    _check(q0)
    C, D = q0.shape
    dev = q0.device
    L, Ls = _per_chain_steps(num_integration_steps, C, dev, max_num_integration_steps)
    eps = float(step_size)
    l0 = state.logdensity.reshape(C).to(torch.float64)
    g0 = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
    z, u_acc = _noise(rng_key, C, D, dev)
    G, dG = _metric(metric_fn, q0)
    p0, _, _, logdet0, pu0 = _field(G, None, z=z)
    H0 = -l0 + _kinetic(logdet0, pu0, D)
    q, p, l, g = q0, p0, l0, g0
    infos = []
    for st in range(L):
        qn, pn, info = _implicit_midpoint_step(q, p, eps, logdensity_fn, metric_fn, solver)
        if Ls is None:
            q, p = qn, pn
        else:
            act = st < Ls
            q, p = torch.where(act[:, None], qn, q), torch.where(act[:, None], pn, p)
            info = (info[0] & act, torch.where(act, info[1], torch.zeros_like(info[1])),
                    torch.where(act, info[2], torch.zeros_like(info[2])))
        infos.append(info)
    if L > 0:
        l, g = _value_and_grad(logdensity_fn, q)
        G, dG = _metric(metric_fn, q)
        if Ls is not None:  # chains that took no step keep the state's own values
            moved = Ls > 0
            l, g = torch.where(moved, l, l0), torch.where(moved[:, None], g, g0)
    _, v, _, logdet1, pu1 = _field(G, None, p=p)
    H1 = -l + _kinetic(logdet1, pu1, D)
    delta = H0 - H1
    delta = torch.where(torch.isnan(delta), torch.full_like(delta, -float("inf")), delta)
    p_accept = torch.clamp(torch.exp(delta), max=1.0)
    accept = u_acc < p_accept
    end = IntegratorState(q, -p, -v, l, g)  # flipped (geomjax/rmhmc/rmhmc.py:443-462)
    proposal = Proposal(end, H1, delta, torch.minimum(delta, torch.zeros_like(delta)))
    a1 = accept[:, None]
    new_state = RMHMCState(torch.where(a1, q, q0), torch.where(accept, l, l0), torch.where(a1, g, g0))
    info = RMHMCInfo(p0, p_accept, accept, -delta > divergence_threshold, H1, proposal, L if Ls is None else Ls)
    return new_state, info, infos


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples, solver, *,
           divergence_threshold=1000, store_samples=True, samples_out=None, max_num_integration_steps=None):
    """``num_samples`` transitions (a Python-driven loop; capture it with
    ``GraphSampler`` to remove the host dispatch)."""
    from .rmhmc import RMHMCState
    from .sampling import SampleResult

    q = state.position
    C, D = q.shape
    T = int(num_samples)
    if T < 1:
        raise ValueError("num_samples must be >= 1")
    dev = q.device
    samples = samples_out if samples_out is not None else (
        torch.zeros(T, C, D, dtype=torch.float64, device=dev) if store_samples else None)
    acc = torch.empty(T, C, dtype=torch.float32, device=dev)
    n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
    fev = torch.zeros(C, dtype=torch.int64, device=dev)
    s = RMHMCState(*state)
    for t in range(T):
        key = PRNGKey(rng_key.seed, rng_key.counter + t) if isinstance(rng_key, PRNGKey) else \
            HostNoise(rng_key.z[t], rng_key.u[t])
        steps = num_integration_steps[t] if isinstance(num_integration_steps, torch.Tensor) else num_integration_steps
        s, info, fp = transition(key, s, logdensity_fn, step_size, metric_fn, steps, solver, divergence_threshold,
                                 max_num_integration_steps)
        if samples is not None:
            samples[t].copy_(s.position)
        acc[t] = info.acceptance_rate.float()
        n_acc += info.is_accepted.int()
        # per step: the first evaluation, the n iterations, the explicit half step
        for (_, _, n) in fp:
            fev += n
        fev += 2 * (info.num_integration_steps if isinstance(info.num_integration_steps, torch.Tensor)
                    else torch.full_like(fev, info.num_integration_steps))
    return SampleResult(s, samples, acc, n_acc, fev)
