"""LMC with a general dense user metric (geomjax/lmcmc with any ``metric_fn``).

The reference's lmcmc accepts any ``metric_fn`` and differentiates it with
``jax.jacfwd`` (geomjax/lmcmc/metrics.py:158-189).  Here the user's functions
(``logdensity_fn`` and ``metric_fn``, written with torch operations that
``torch.func`` can transform, on the device of their input) are evaluated and
differentiated by ``torch.func`` on the GPU, batched over chains; the
integrator's linear algebra runs in the HIP kernels of ``dense_metric.hip``:
the D^3 Christoffel contractions, the LU factorisations with partial pivoting
of G +- (eps/2) Omega~, the solve, the log-determinants and the volume change
(``gjx_dense_lmc_half_step``), and the Cholesky velocity draw, log det G and
grad log det G = tr(G^-1 dG) (``gjx_dense_metric_position``).  float64,
D <= 64.

With a ``PRNGKey`` the normals and uniforms are drawn by torch's device
generator seeded from (seed, counter); ``HostNoise(z [C, D], u [C])`` feeds
them explicitly.
"""
import torch
from torch.func import grad_and_value, jacfwd, vmap

from . import _lib
from .proposal import Proposal
from .random import HostNoise, PRNGKey

__all__ = ["is_dense_metric_fn", "init", "transition", "sample"]


def is_dense_metric_fn(logdensity_fn, metric_fn):
    """True for plain callables (user functions) -- including ``MongeMetric``
    paired with a user log density, whose metric is then built by autodiff --
    False for this package's model descriptors, which have closed-form kernels."""
    from .metrics import FisherMetric, MongeMetric
    from .models import LogisticBatch, LogisticRegression, NealFunnel

    if isinstance(logdensity_fn, (LogisticRegression, LogisticBatch, NealFunnel)):
        return False
    if isinstance(metric_fn, FisherMetric):
        return False
    if isinstance(metric_fn, MongeMetric):
        return callable(logdensity_fn)
    return callable(metric_fn) and callable(logdensity_fn)


def resolve_metric_fn(logdensity_fn, metric_fn):
    """The metric as a function of the position: ``MongeMetric(alpha2)`` with a
    user log density becomes I + alpha2 grad l grad l^T (the Monge metric the
    reference's examples pass as metric_fn, examples/funnel/main.py:34-39)."""
    from .metrics import MongeMetric

    if not isinstance(metric_fn, MongeMetric):
        return metric_fn
    alpha2 = metric_fn.alpha2
    g = torch.func.grad(logdensity_fn)

    def monge(q):
        gq = g(q)
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + alpha2 * torch.outer(gq, gq)

    return monge


def _check(position):
    if not position.is_cuda or position.dtype != torch.float64:
        raise ValueError("dense-metric sampler states must be float64 tensors on a HIP device")
    if position.dim() != 2:
        raise ValueError("the dense-metric samplers run a batch of chains: position must be [C, D]")
    if position.shape[1] > 64:
        raise NotImplementedError(f"dense-metric kernels take D <= 64 (got {position.shape[1]})")


def _as_kernel_input(t, q, shape, what):
    """A user function's output as the float64 kernels read it: the exact
    ``shape``, float64 and on ``q``'s device (a float32 result is widened, a
    host result moved), contiguous.  Anything that is not a real floating
    tensor of that shape raises before a launch."""
    if not torch.is_tensor(t):
        raise TypeError(f"{what} must return a torch tensor (got {type(t).__name__})")
    if not (t.is_floating_point() and not t.is_complex()):
        raise TypeError(f"{what} must return a real floating-point tensor (got {t.dtype})")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} returned shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t.to(device=q.device, dtype=torch.float64).contiguous()


def _value_and_grad(logdensity_fn, q):
    g, l = vmap(grad_and_value(logdensity_fn))(q)
    C, D = q.shape
    return (_as_kernel_input(l, q, (C,), "logdensity_fn (a scalar per chain)"),
            _as_kernel_input(g, q, (C, D), "the gradient of logdensity_fn"))


def _metric(metric_fn, q):
    """G [C, D, D] and dG [C, D, D, D] (dG_ab / dq_i at [c, a, b, i]).  A metric_fn
    returning a vector is a diagonal metric (the reference's ndim == 1 branch,
    geomjax/rmhmc/metrics.py:45-65): it is embedded as diag(m), dG_aa/dq_i = dm_a/dq_i.
    The results are float64 on q's device whatever the function returned
    (float32 results are widened, host results moved)."""
    C, D = q.shape
    G = vmap(metric_fn)(q)
    J = vmap(jacfwd(metric_fn))(q)
    if torch.is_tensor(G) and G.dim() == 2:  # [C, D]: diagonal
        G = _as_kernel_input(G, q, (C, D), "metric_fn (diagonal)")
        J = _as_kernel_input(J, q, (C, D, D), "the Jacobian of metric_fn (diagonal)")
        G = torch.diag_embed(G)
        J = torch.diag_embed(J.transpose(1, 2)).permute(0, 2, 3, 1)  # [C, i, a, a] -> [C, a, b, i]
    elif torch.is_tensor(G) and G.dim() == 3:
        G = _as_kernel_input(G, q, (C, D, D), "metric_fn")
        J = _as_kernel_input(J, q, (C, D, D, D), "the Jacobian of metric_fn")
    else:
        got = tuple(G.shape[1:]) if torch.is_tensor(G) else type(G).__name__
        raise ValueError(f"metric_fn must return a [D] or [D, D] tensor (got {got})")
    return G.contiguous(), J.contiguous()


def _io(G, dG, **kw):
    C, D = G.shape[0], G.shape[1]
    ptrs = {k: _lib.ptr(v) for k, v in kw.items() if k != "step_size"}
    return _lib.DenseMetricIO(num_chains=C, dim=D, step_size=float(kw.get("step_size", 0.0)), metric=_lib.ptr(G),
                              metric_jacobian=_lib.ptr(dG), **ptrs)


def _position(G, dG, z=None, velocity=None):
    C, D = G.shape[0], G.shape[1]
    dev = G.device
    logdet = torch.empty(C, dtype=torch.float64, device=dev)
    gld = torch.empty(C, D, dtype=torch.float64, device=dev)
    v0 = torch.empty(C, D, dtype=torch.float64, device=dev) if z is not None else None
    vgv = torch.empty(C, dtype=torch.float64, device=dev) if velocity is not None else None
    io = _io(G, dG, z=z, velocity=velocity, velocity_out=v0, logdet=logdet, grad_logdet=gld, vgv=vgv)
    _lib.check(_lib.load().gjx_dense_metric_position(io, _lib.stream_handle(dev)), "gjx_dense_metric_position")
    return logdet, gld, v0, vgv


def _half_step(G, dG, v, g, gld, eps, J):
    """One half step (``eps`` a float or a [C] tensor of signed step sizes);
    returns the new velocity, J updated in place."""
    out = torch.empty_like(v)
    if torch.is_tensor(eps):
        io = _io(G, dG, step_size=0.0, step_sizes=eps.to(torch.float64).contiguous(), velocity=v,
                 logdensity_grad=g, grad_logdet_in=gld, velocity_out=out, volume_adjustment=J)
    else:
        io = _io(G, dG, step_size=eps, velocity=v, logdensity_grad=g, grad_logdet_in=gld, velocity_out=out,
                 volume_adjustment=J)
    _lib.check(_lib.load().gjx_dense_lmc_half_step(io, _lib.stream_handle(v.device)), "gjx_dense_lmc_half_step")
    return out


def lmc_step(state, eps, logdensity_fn, metric_fn):
    """One lan_integrator step (integrators.py:93-142) of every chain from the tuple
    (q, momentum, velocity, l, g, J), ``eps`` a [C] tensor of signed step sizes;
    returns the new tuple and its energy -l - log det G / 2 + v^T G v / 2 - J
    (metrics.py:209-221)."""
    q, _, v, l, g, J = state
    G, dG = _metric(metric_fn, q)
    _, gld, _, _ = _position(G, dG)
    J = J.clone()
    v = _half_step(G, dG, v, g, gld, eps, J)
    q = q + eps[:, None] * v
    l, g = _value_and_grad(logdensity_fn, q)
    G, dG = _metric(metric_fn, q)
    logdet, gld, _, _ = _position(G, dG)
    v = _half_step(G, dG, v, g, gld, eps, J)
    _, _, _, vgv = _position(G, dG, velocity=v)
    momentum = torch.bmm(G, v[:, :, None])[:, :, 0]
    return (q, momentum, v, l, g, J), -l + (-0.5 * logdet + 0.5 * vgv) - J


def init(position, logdensity_fn):
    """geomjax/lmcmc/lmc.py:98-101 for a user log density."""
    from .lmcmc import LMCState

    single = position.dim() == 1
    q = (position[None] if single else position).contiguous()
    _check(q)
    _lib.load()  # the sampler's kernels are native: fail here, not at the first step, without them
    l, g = _value_and_grad(logdensity_fn, q)
    if single:
        return LMCState(position, l[0], g[0], torch.zeros_like(l[0]))
    return LMCState(position, l, g, torch.zeros_like(l))


def as_batch(state):
    """A single-chain state (position [D]) as a batch of one chain, and whether
    it was single: every tensor field gains a leading axis of 1."""
    if state.position.dim() != 1:
        return state, False
    return type(state)(*((t[None] if torch.is_tensor(t) else t) for t in state)), True


def unbatch(x):
    """The inverse of ``as_batch`` on a transition's outputs: the leading chain
    axis of every tensor in (nested) tuples, lists and dicts is dropped."""
    if torch.is_tensor(x):
        return x[0] if x.dim() >= 1 else x
    if isinstance(x, tuple) and hasattr(x, "_fields"):
        return type(x)(*(unbatch(t) for t in x))
    if isinstance(x, (tuple, list)):
        return type(x)(unbatch(t) for t in x)
    if isinstance(x, dict):
        return {k: unbatch(v) for k, v in x.items()}
    return x


def _noise(rng_key, C, D, dev):
    if isinstance(rng_key, HostNoise):
        return (rng_key.z.reshape(C, D).to(device=dev, dtype=torch.float64).contiguous(),
                rng_key.u.reshape(C).to(device=dev, dtype=torch.float64).contiguous())
    if isinstance(rng_key, PRNGKey):
        gen = torch.Generator(device=dev)
        gen.manual_seed((int(rng_key.seed) * 0x9E3779B1 + int(rng_key.counter)) & 0x7FFFFFFFFFFFFFFF)
        return (torch.randn(C, D, generator=gen, dtype=torch.float64, device=dev),
                torch.rand(C, generator=gen, dtype=torch.float64, device=dev))
    raise TypeError("rng_key must be geomjax_amd.random.HostNoise or PRNGKey")


def transition(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps,
               divergence_threshold=1000.0, max_num_integration_steps=None):
    """One LMC transition (geomjax/lmcmc/lmc.py:135-180, 411-475) of every chain;
    ``num_integration_steps`` an int or a [C] tensor of per-chain counts (a chain
    whose count is reached keeps its state while the others step on)."""
    from .lmcmc import IntegratorState, LMCInfo, LMCState
    from .rmhmc_dense import _per_chain_steps

    metric_fn = resolve_metric_fn(logdensity_fn, metric_fn)

    q0 = state.position.contiguous()
    _check(q0)
    C, D = q0.shape
    dev = q0.device
    L, Ls = _per_chain_steps(num_integration_steps, C, dev, max_num_integration_steps)
    eps = float(step_size)
    l0 = state.logdensity.reshape(C).to(torch.float64)
    g0 = state.logdensity_grad.reshape(C, D).to(torch.float64).contiguous()
    J0 = torch.as_tensor(state.volume_adjustment, dtype=torch.float64, device=dev).expand(C).clone()
    z, u = _noise(rng_key, C, D, dev)

    G, dG = _metric(metric_fn, q0)
    logdet, gld, v0, _ = _position(G, dG, z=z)
    _, _, _, vgv0 = _position(G, dG, velocity=v0)
    H0 = -l0 + (-0.5 * logdet + 0.5 * vgv0) - J0

    q, v, l, g, J = q0, v0, l0, g0, J0.clone()
    for st in range(L):
        Jn = J.clone()
        vn = _half_step(G, dG, v, g, gld, eps, Jn)
        qn = q + eps * vn
        ln, gn = _value_and_grad(logdensity_fn, qn)
        Gn, dGn = _metric(metric_fn, qn)
        logdetn, gldn, _, _ = _position(Gn, dGn)
        vn = _half_step(Gn, dGn, vn, gn, gldn, eps, Jn)
        if Ls is None:
            q, v, l, g, J, G, dG, logdet, gld = qn, vn, ln, gn, Jn, Gn, dGn, logdetn, gldn
        else:
            a = st < Ls
            a1, a2, a3 = a[:, None], a[:, None, None], a[:, None, None, None]
            q, v, g, gld = (torch.where(a1, x, y) for x, y in ((qn, q), (vn, v), (gn, g), (gldn, gld)))
            l, J, logdet = (torch.where(a, x, y) for x, y in ((ln, l), (Jn, J), (logdetn, logdet)))
            G, dG = torch.where(a2, Gn, G).contiguous(), torch.where(a3, dGn, dG).contiguous()
    _, _, _, vgv1 = _position(G, dG, velocity=v)
    H1 = -l + (-0.5 * logdet + 0.5 * vgv1) - J
    delta = H0 - H1
    delta = torch.where(torch.isnan(delta), torch.full_like(delta, -float("inf")), delta)
    p_accept = torch.clamp(torch.exp(delta), max=1.0)
    accept = u < p_accept
    momentum = torch.bmm(G, v[:, :, None])[:, :, 0]
    end = IntegratorState(q, -momentum, -v, l, g, J)  # flipped (geomjax/lmcmc/lmc.py:478-499)
    proposal = Proposal(end, H1, delta, torch.minimum(delta, torch.zeros_like(delta)))
    a2 = accept[:, None]
    new_state = LMCState(torch.where(a2, q, q0), torch.where(accept, l, l0), torch.where(a2, g, g0),
                         torch.where(accept, J, J0))
    info = LMCInfo(v0, p_accept, accept, -delta > divergence_threshold, H1, proposal, L if Ls is None else Ls)
    return new_state, info


def sample(rng_key, state, logdensity_fn, step_size, metric_fn, num_integration_steps, num_samples, *,
           divergence_threshold=1000, store_samples=True, samples_out=None, max_num_integration_steps=None):
    """``num_samples`` transitions (one Python-driven loop of kernel launches)."""
    from .lmcmc import LMCState
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
    s = LMCState(*state)
    for t in range(T):
        key = rng_key if isinstance(rng_key, PRNGKey) else HostNoise(rng_key.z[t], rng_key.u[t])
        if isinstance(key, PRNGKey):
            key = PRNGKey(key.seed, key.counter + t)
        steps = num_integration_steps[t] if isinstance(num_integration_steps, torch.Tensor) else num_integration_steps
        s, info = transition(key, s, logdensity_fn, step_size, metric_fn, steps, divergence_threshold,
                             max_num_integration_steps)
        if samples is not None:
            samples[t].copy_(s.position)
        acc[t] = info.acceptance_rate.float()
        n_acc += info.is_accepted.int()
        n = info.num_integration_steps
        fev += n.to(torch.int64) if isinstance(n, torch.Tensor) else n
    return SampleResult(s, samples, acc, n_acc, fev)
