"""Target densities served by the HIP kernels.

The reference takes arbitrary JAX callables ``logdensity_fn`` and
``metric_fn`` and differentiates them with ``jax.grad``/``jvp``/``jacfwd``.
A GPU kernel cannot trace a Python callable, so the boundary accepts a closed
set of model descriptors whose log density, gradient and Hessian-vector
products are written in closed form inside the kernels.
"""
import ctypes

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


class LogisticRegression:
    """Bayesian logistic regression: y_n ~ Bernoulli(sigmoid(x_n . q)), q ~ N(0, prior_var I).

    The reference leaves the model user-supplied; the log density is
    sum_n [y_n eta_n - log(1 + e^eta_n)] - |q|^2 / (2 prior_var)
    - (D/2) log(2 pi prior_var) with eta = X q.  ``X`` [N, D] and ``y`` [N]
    are copied to the device once (float64).
    """

    kind = "logistic"

    def __init__(self, X, y, prior_var=100.0, device="cuda:0"):
        X = torch.as_tensor(X, dtype=torch.float64)
        y = torch.as_tensor(y, dtype=torch.float64)
        if X.dim() != 2 or y.shape != (X.shape[0],):
            raise ValueError(f"X must be [N, D] and y [N] (got {tuple(X.shape)} and {tuple(y.shape)})")
        if not prior_var > 0:
            raise ValueError("prior_var must be > 0")
        self.N, self.D = X.shape
        self.prior_var = float(prior_var)
        self.X = X.to(device).contiguous()
        self.y = y.to(device).contiguous()
        self._table = _posterior_table([self], None, self.X.device)

    @property
    def dim(self):
        return self.D

    def data(self, chain_posterior=None, *, dtype=torch.float64, num_chains=None):
        return self._table.struct(chain_posterior, dtype=dtype, num_chains=num_chains)

    def value_and_grad(self, position):
        return logistic_value_and_grad(self, position)

    def __repr__(self):
        return f"LogisticRegression(N={self.N}, D={self.D}, prior_var={self.prior_var})"


class LogisticBatch:
    """Many independent logistic-regression posteriors sampled in one launch.

    ``models`` are LogisticRegression descriptors (same prior variance);
    ``chains_per_posterior`` chains are assigned to each, in order.  States are
    [C, Dmax] with Dmax = max D_p; coordinates past a chain's own D_p are zero
    and stay zero.
    """

    kind = "logistic"

    def __init__(self, models, chains_per_posterior):
        if not models:
            raise ValueError("LogisticBatch needs at least one model")
        pv = {m.prior_var for m in models}
        if len(pv) != 1:
            raise ValueError("all posteriors of a batch must share prior_var")
        self.models = list(models)
        self.prior_var = pv.pop()
        counts = [int(chains_per_posterior)] * len(models) if isinstance(chains_per_posterior, int) \
            else [int(c) for c in chains_per_posterior]
        if len(counts) != len(models) or min(counts) < 1:
            raise ValueError("chains_per_posterior must give >= 1 chain per model")
        self.counts = counts
        self.D = max(m.D for m in models)
        self.N = max(m.N for m in models)
        dev = models[0].X.device
        self.chain_posterior = torch.repeat_interleave(torch.arange(len(models), dtype=torch.int32),
                                                       torch.tensor(counts)).to(dev)
        self._table = _posterior_table(self.models, self.chain_posterior, dev)

    @property
    def dim(self):
        return self.D

    @property
    def num_chains(self):
        return int(sum(self.counts))

    def data(self, chain_posterior=None, *, dtype=torch.float64, num_chains=None):
        return self._table.struct(self.chain_posterior, dtype=dtype, num_chains=num_chains)

    def value_and_grad(self, position):
        return logistic_value_and_grad(self, position)


class _posterior_table:
    def __init__(self, models, chain_posterior, dev):
        self.x = torch.cat([m.X.reshape(-1) for m in models]).to(dev)
        self.y = torch.cat([m.y for m in models]).to(dev)
        xoff, yoff, ox, oy = [], [], 0, 0
        for m in models:
            xoff.append(ox)
            yoff.append(oy)
            ox += m.N * m.D
            oy += m.N
        self.N = torch.tensor([m.N for m in models], dtype=torch.int32, device=dev)
        self.Dp = torch.tensor([m.D for m in models], dtype=torch.int32, device=dev)
        self.xoff = torch.tensor(xoff, dtype=torch.int64, device=dev)
        self.yoff = torch.tensor(yoff, dtype=torch.int64, device=dev)
        self.P = len(models)
        self.max_rows = max(m.N for m in models)
        self.max_features = max(m.D for m in models)
        self._host_N = [m.N for m in models]
        self._host_D = [m.D for m in models]
        self._packed = None     # float32 fragment table (built on first float32 use)
        self._workspace = None  # per-chain scratch of the float32 path
        self._retired = []      # outgrown scratch buffers: a captured HIP graph may still point at them
        self._orders = {}       # num_chains -> dispatch order of the float32 kernels

    def _base(self, chain_posterior):
        return _lib.LogisticData(num_posteriors=self.P, max_rows=self.max_rows, max_features=self.max_features,
                                 reserved=0, num_rows=self.N.data_ptr(), num_features=self.Dp.data_ptr(),
                                 x_offset=self.xoff.data_ptr(), y_offset=self.yoff.data_ptr(),
                                 chain_posterior=_lib.ptr(chain_posterior), x=self.x.data_ptr(), y=self.y.data_ptr(),
                                 packed=None, workspace=None, workspace_bytes=0)

    def packed(self):
        """The device table of X in the float32 kernel's fragment layouts (gjx_logistic_pack)."""
        if self._packed is None:
            lib = _lib.load()
            n = (ctypes.c_int32 * self.P)(*self._host_N)
            d = (ctypes.c_int32 * self.P)(*self._host_D)
            nbytes = ctypes.c_uint64()
            _lib.check(lib.gjx_logistic_pack_size(self.P, n, d, ctypes.byref(nbytes)), "gjx_logistic_pack_size")
            buf = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.x.device)
            rc = lib.gjx_logistic_pack(self._base(None), n, d, buf.data_ptr(), nbytes.value,
                                       _lib.stream_handle(self.x.device))
            _lib.check(rc, "gjx_logistic_pack")
            self._packed = buf
        return self._packed

    def workspace(self, num_chains):
        need = num_chains * 5 * ((self.max_rows + 63) // 64 * 64) * 4
        if self._workspace is None or self._workspace.numel() < need:
            if self._workspace is not None:
                self._retired.append(self._workspace)
            self._workspace = torch.empty(need, dtype=torch.uint8, device=self.x.device)
        return self._workspace

    def dispatch_order(self, chain_posterior, num_chains):
        """Order in which the float32 kernels start the chains of a batch.

        The kernels deal dispatch slots to the 8 XCDs in contiguous ranges of
        ceil(C/8) slots.  Posteriors are taken by decreasing cost per chain
        (N times the square of the padded width) and each goes whole to the XCD
        with the least cost so far that has room (spilling into the next when
        none has room for all its chains); inside an XCD the chains keep that
        order, so the costliest start first and the cheap ones fill the launch's
        tail, and a posterior's X stays in few L2 caches.  None for a single
        posterior (every chain costs the same)."""
        if chain_posterior is None or self.P == 1:
            return None
        C = int(num_chains)
        if C in self._orders:
            return self._orders[C]
        cp = chain_posterior.detach().to("cpu", torch.int64)
        if cp.numel() != C:
            return None
        pad = [32 if d <= 32 else (64 if d <= 64 else 128) for d in self._host_D]
        cost = [n * w * w for n, w in zip(self._host_N, pad)]
        per = (C + 7) // 8
        cap = [max(0, min(per, C - x * per)) for x in range(8)]
        lists = [[] for _ in range(8)]
        load = [0.0] * 8
        chains_of = [[] for _ in range(self.P)]
        for c, p in enumerate(cp.tolist()):
            if 0 <= p < self.P:
                chains_of[p].append(c)
        for p in sorted(range(self.P), key=lambda q: -cost[q]):
            rest = chains_of[p]
            while rest:
                room = [x for x in range(8) if len(lists[x]) < cap[x]]
                fits = [x for x in room if cap[x] - len(lists[x]) >= len(rest)]
                x = min(fits or room, key=lambda k: load[k])
                take = rest[:cap[x] - len(lists[x])]
                rest = rest[len(take):]
                lists[x].extend(take)
                load[x] += cost[p] * len(take)
        order = [c for x in range(8) for c in lists[x]]
        placed = set(order)
        order += [c for c in range(C) if c not in placed]  # chains of an invalid posterior index
        t = torch.tensor(order, dtype=torch.int32, device=self.x.device)
        self._orders[C] = t
        return t

    def struct(self, chain_posterior, dtype=torch.float64, num_chains=None):
        st = self._base(chain_posterior)
        if dtype == torch.float32:
            if num_chains is None:
                raise ValueError("the float32 path needs num_chains to size its workspace")
            ws = self.workspace(num_chains)
            st.packed = self.packed().data_ptr()
            st.workspace = ws.data_ptr()
            st.workspace_bytes = ws.numel()
            st.dispatch_order = _lib.ptr(self.dispatch_order(chain_posterior, num_chains))
        return st


def logistic_params(model, C, L=0, T=1, step_size=0.0, max_iters=0, convergence_tol=1e-6, divergence_tol=1e10,
                    divergence_threshold=1000.0, seed=0, counter=0, dtype=torch.float64, waves_per_chain=0):
    code = _lib.GJX_F32 if dtype == torch.float32 else _lib.GJX_F64
    return _lib.LogisticParams(num_chains=C, dim=model.dim, dtype=code, num_integration_steps=int(L),
                               max_iters=int(max_iters), num_transitions=int(T), flags=0, waves_per_chain=int(waves_per_chain),
                               step_size=float(step_size), prior_var=model.prior_var,
                               convergence_tol=float(convergence_tol), divergence_tol=float(divergence_tol),
                               divergence_threshold=float(divergence_threshold), seed=int(seed),
                               counter=int(counter))


def logistic_value_and_grad(model, position):
    """Log density (float64) and gradient (in the position's dtype) of C positions.

    float32 positions are evaluated by the float64 kernel on a widened copy
    (this is the reference's init, not the sampling loop)."""
    if not position.is_cuda:
        raise ValueError("positions must be on a HIP device (cuda:N)")
    if position.dtype not in (torch.float64, torch.float32):
        raise TypeError("logistic positions must be float64 or float32")
    if position.dtype == torch.float32:
        logp, grad = logistic_value_and_grad(model, position.double())
        return logp, grad.float()
    pos = position.contiguous()
    squeeze = pos.dim() == 1
    if squeeze:
        pos = pos[None]
    C, D = pos.shape
    if D != model.dim:
        raise ValueError(f"position has {D} coordinates, the model has {model.dim}")
    if isinstance(model, LogisticBatch) and C != model.num_chains:
        raise ValueError(f"the batch has {model.num_chains} chains, got {C} positions")
    logp = torch.empty(C, dtype=torch.float64, device=pos.device)
    grad = torch.zeros_like(pos)
    rc = _lib.load().gjx_logistic_value_and_grad(logistic_params(model, C), model.data(), pos.data_ptr(),
                                                 logp.data_ptr(), grad.data_ptr(), _lib.stream_handle(pos.device))
    _lib.check(rc, "gjx_logistic_value_and_grad")
    if squeeze:
        return logp[0], grad[0]
    return logp, grad
