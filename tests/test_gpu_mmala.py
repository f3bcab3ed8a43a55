"""GPU parity: manifold MALA with the Fisher metric on Bayesian logistic
regression (float64) against the oracle (oracle/mmala.py: jacfwd Gamma,
dense inverse, Cholesky of the covariance, multivariate-normal log densities).

Bars: positions, log densities, gradients, metrics, Gamma, weights and
acceptance statistics 1e-10 norm-wise relative; accept decisions and
divergence flags identical."""
import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.mmala import DiffusionState, build_kernel
from geomjax_amd.mmala import init as mm_init
from geomjax_amd.synthetic import make_logistic
from oracle import mmala as omm
from oracle import models as om
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _cpu(s):
    return DiffusionState(*(t.cpu() for t in s))


@pytest.mark.parametrize("N,D,C,eps,seed", [(120, 6, 3, 0.05, 1), (256, 64, 2, 0.01, 2)])
def test_init_matches_oracle(N, D, C, eps, seed):
    print(f"case mmala init N={N} D={D} C={C}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    model, Mo = gj.LogisticRegression(X, y, 100.0), om.LogisticRegression(X, y, 100.0)
    q0 = 0.3 * seeded_normal(10 + seed, C, D)
    s = _cpu(mm_init(q0.to(DEV), model, gj.FisherMetric()))
    so = omm.init(q0, Mo.logp, Mo.fisher_metric_fn)
    for f in ("logdensity", "logdensity_grad", "metric", "Gamma"):
        assert rel_err(getattr(s, f), getattr(so, f)) < 1e-10, f


@pytest.mark.parametrize("N,D,C,eps,T,faithful,seed", [
    (120, 6, 4, 0.4, 3, True, 3),
    (120, 6, 4, 0.4, 3, False, 4),
    (300, 13, 3, 0.2, 2, True, 5),
    (256, 64, 2, 0.05, 2, False, 6),    # the kernel's full feature capacity
    (120, 6, 4, 2.0, 2, True, 7),       # a step far too large: proposals far out, weights huge
])
def test_step_matches_oracle(N, D, C, eps, T, faithful, seed):
    print(f"case mmala step N={N} D={D} C={C} eps={eps} T={T} faithful={faithful}", flush=True)
    X, y, _ = make_logistic(N, D, seed)
    model, Mo = gj.LogisticRegression(X, y, 100.0), om.LogisticRegression(X, y, 100.0)
    q0 = 0.3 * seeded_normal(20 + seed, C, D)
    so = omm.init(q0, Mo.logp, Mo.fisher_metric_fn)
    s = mm_init(q0.to(DEV), model, gj.FisherMetric())
    kern = build_kernel(reference_faithful=faithful)
    okern = omm.build_kernel(reference_faithful=faithful)
    for t in range(T):
        z, u = seeded_normal(30 + 10 * seed + t, C, D), seeded_uniform(40 + 10 * seed + t, C)
        s1, info = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, model, eps, gj.FisherMetric())
        so1, oi = okern((z, u), so, Mo.logp, eps, Mo.fisher_metric_fn)
        torch.cuda.synchronize()
        print(f"  t={t}: accepted {oi.is_accepted.tolist()}, weight {[round(float(w), 4) for w in oi.weight]}",
              flush=True)
        assert info.is_accepted.cpu().tolist() == oi.is_accepted.tolist()
        assert info.is_divergent.cpu().tolist() == oi.is_divergent.tolist()
        wscale = max(1.0, oi.weight.abs().max().item())
        assert (info.weight.cpu() - oi.weight).abs().max().item() < 1e-10 * wscale
        assert rel_err(info.acceptance_rate, oi.acceptance_rate) < 1e-10
        for f in DiffusionState._fields:
            assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, ("state", f)
            assert rel_err(getattr(info.proposal, f), getattr(oi.proposal, f)) < 1e-10, ("proposal", f)
        so, s = so1, s1


def test_sample_kernel_matches_step_kernel():
    X, y, _ = make_logistic(200, 8, 8)
    model = gj.LogisticRegression(X, y, 100.0)
    C, D, T, eps = 16, 8, 5, 0.05
    z, u = seeded_normal(50, T, C, D), seeded_uniform(51, T, C)
    s0 = mm_init((0.2 * seeded_normal(52, C, D)).to(DEV), model, gj.FisherMetric())
    for faithful in (True, False):
        res = gj.mmala(model, eps, gj.FisherMetric(), reference_faithful=faithful).sample_fn(
            gr.HostNoise(z.to(DEV), u.to(DEV)), s0, T)
        kern = build_kernel(reference_faithful=faithful)
        s, n_acc = s0, torch.zeros(C, dtype=torch.int32)
        for t in range(T):
            s, info = kern(gr.HostNoise(z[t].to(DEV), u[t].to(DEV)), s, model, eps, gj.FisherMetric())
            assert rel_err(res.samples[t], s.position) <= 1e-12
            assert rel_err(res.acceptance_rate[t].double(), info.acceptance_rate.float().double()) <= 1e-6
            n_acc += info.is_accepted.cpu().int()
        torch.cuda.synchronize()
        assert torch.equal(res.num_accepted.cpu(), n_acc)
        for f in DiffusionState._fields:
            assert rel_err(getattr(res.state, f), getattr(s, f)) <= 1e-12, f


def test_batch_equals_each_posterior_alone():
    Xa, ya, _ = make_logistic(150, 5, 9)
    Xb, yb, _ = make_logistic(90, 9, 10)
    ma, mb = gj.LogisticRegression(Xa, ya, 100.0), gj.LogisticRegression(Xb, yb, 100.0)
    batch = gj.LogisticBatch([ma, mb], [3, 2])
    C, Dm = 5, 9
    q = torch.zeros(C, Dm, dtype=torch.float64)
    q[:3, :5] = 0.3 * seeded_normal(61, 3, 5)
    q[3:, :9] = 0.3 * seeded_normal(62, 2, 9)
    z = torch.zeros(C, Dm, dtype=torch.float64)
    z[:3, :5] = seeded_normal(63, 3, 5)
    z[3:, :9] = seeded_normal(64, 2, 9)
    u = seeded_uniform(65, C)
    kern = build_kernel()
    nb, ib = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), mm_init(q.to(DEV), batch), batch, 0.05, gj.FisherMetric())
    for model, rows, D in ((ma, slice(0, 3), 5), (mb, slice(3, 5), 9)):
        s = mm_init(q[rows, :D].contiguous().to(DEV), model)
        n1, i1 = kern(gr.HostNoise(z[rows, :D].contiguous().to(DEV), u[rows].to(DEV)), s, model, 0.05,
                      gj.FisherMetric())
        torch.cuda.synchronize()
        assert torch.equal(ib.is_accepted[rows], i1.is_accepted)
        assert rel_err(nb.position[rows, :D], n1.position) <= 1e-12
        assert rel_err(nb.metric[rows, :D, :D], n1.metric) <= 1e-12
        assert rel_err(nb.Gamma[rows, :D], n1.Gamma) <= 1e-12
        assert torch.count_nonzero(nb.position[rows, D:]) == 0
        assert torch.count_nonzero(nb.metric[rows, D:, :]) == 0 and torch.count_nonzero(nb.metric[rows, :, D:]) == 0


def test_exact_mode_agrees_with_rmhmc_on_the_posterior_moments():
    """The corrected test (reference_faithful=False) targets the posterior: after
    burn-in its chains' mean and variance of each coefficient agree with those of
    the float64 RMHMC kernel's chains (4096 chains each, 6 standard errors of the
    difference of two means; variances within 15 %)."""
    import functools

    from geomjax_amd.rmhmc import implicit_midpoint

    X, y, _ = make_logistic(400, 4, 11)
    model = gj.LogisticRegression(X, y, 100.0)
    C = 4096
    q0 = (0.1 * torch.randn(C, 4, generator=torch.Generator().manual_seed(12), dtype=torch.float64)).to(DEV)
    mm = gj.mmala(model, 0.3, gj.FisherMetric(), reference_faithful=False)
    rm = gj.rmhmc(model, 0.2, gj.FisherMetric(), 8, integrator=functools.partial(implicit_midpoint, max_iters=8))
    a = mm.sample_fn(gr.key(13), mm.init(q0), 300, store_samples=False)
    b = rm.sample_fn(gr.key(14), rm.init(q0), 60, store_samples=False)
    torch.cuda.synchronize()
    qa, qb = a.state.position.cpu(), b.state.position.cpu()
    acc = float(a.acceptance_rate[-50:].mean())
    print(f"mmala mean {qa.mean(0).tolist()} var {qa.var(0).tolist()} acceptance {acc:.3f}")
    print(f"rmhmc mean {qb.mean(0).tolist()} var {qb.var(0).tolist()}")
    assert acc > 0.3
    se = (qa.var(0) / C + qb.var(0) / C).sqrt()
    assert ((qa.mean(0) - qb.mean(0)).abs() < 6 * se).all()
    assert ((qa.var(0) / qb.var(0) - 1).abs() < 0.15).all()


def _funnel_logp(q):
    import math

    x, v = q[:-1], q[-1]
    return (-0.5 * v * v / 9.0 - 0.5 * math.log(2 * math.pi * 9.0)
            + torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi)))


def _monge(q):
    g = torch.func.grad(_funnel_logp)(q)
    return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + 0.5 * torch.outer(g, g)


@pytest.mark.parametrize("faithful,D", [(True, 8), (False, 8),
                                         (True, 32), (True, 33),   # the two LDS layouts' boundary
                                         (False, 64)])             # the maximum D
def test_user_dense_metric_matches_oracle(faithful, D):
    """A user metric (the funnel's Monge metric by autodiff): the dense-metric
    kernels and torch.func against the oracle's jacfwd of the inverse metric."""
    print(f"case mmala dense funnel-Monge faithful={faithful} D={D}", flush=True)
    C, eps, T = (4, 0.05, 2) if D <= 8 else (2, 0.03, 1)
    q0 = 0.5 * seeded_normal(70, C, D)
    so = omm.init(q0, _funnel_logp, _monge)
    s = mm_init(q0.to(DEV), _funnel_logp, _monge)
    for f in ("logdensity", "logdensity_grad", "metric", "Gamma"):
        assert rel_err(getattr(s, f), getattr(so, f)) < 1e-10, ("init", f)
    kern, okern = build_kernel(reference_faithful=faithful), omm.build_kernel(reference_faithful=faithful)
    for t in range(T):
        z, u = seeded_normal(71 + t, C, D), seeded_uniform(81 + t, C)
        s1, info = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, _funnel_logp, eps, _monge)
        so1, oi = okern((z, u), so, _funnel_logp, eps, _monge)
        torch.cuda.synchronize()
        assert info.is_accepted.cpu().tolist() == oi.is_accepted.tolist()
        assert (info.weight.cpu() - oi.weight).abs().max().item() < 1e-10 * max(1.0, oi.weight.abs().max().item())
        for f in DiffusionState._fields:
            assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, ("state", f)
            assert rel_err(getattr(info.proposal, f), getattr(oi.proposal, f)) < 1e-10, ("proposal", f)
        so, s = so1, s1


def test_fisher_as_user_function_matches_the_closed_form_kernel():
    import math

    X, y, _ = make_logistic(100, 6, 13)
    Xd = torch.as_tensor(X, dtype=torch.float64, device=DEV)
    yd = torch.as_tensor(y, dtype=torch.float64, device=DEV)

    def logp(b):
        eta = Xd @ b
        return (torch.sum(yd * eta - torch.nn.functional.softplus(eta)) - 0.5 * torch.dot(b, b) / 100.0
                - 0.5 * b.shape[0] * math.log(2 * math.pi * 100.0))

    def fisher(b):
        s = torch.sigmoid(Xd @ b)
        return (Xd.T * (s * (1 - s))) @ Xd + torch.eye(b.shape[0], dtype=b.dtype, device=b.device) / 100.0

    C, D = 8, 6
    q0 = (0.3 * seeded_normal(90, C, D)).to(DEV)
    z, u = seeded_normal(91, C, D).to(DEV), seeded_uniform(92, C).to(DEV)
    model = gj.LogisticRegression(X, y, 100.0)
    su, sc = mm_init(q0, logp, fisher), mm_init(q0, model)
    assert rel_err(su.Gamma, sc.Gamma) < 1e-10 and rel_err(su.metric, sc.metric) < 1e-12
    a, ia = build_kernel(reference_faithful=False)(gr.HostNoise(z, u), su, logp, 0.3, fisher)
    b, ib = build_kernel(reference_faithful=False)(gr.HostNoise(z, u), sc, model, 0.3, gj.FisherMetric())
    torch.cuda.synchronize()
    assert torch.equal(ia.is_accepted, ib.is_accepted)
    assert rel_err(ia.weight, ib.weight) < 1e-9
    for f in DiffusionState._fields:
        assert rel_err(getattr(a, f), getattr(b, f)) < 1e-10, f


def test_user_dense_metric_single_chain_equals_the_batched_row():
    C, D, eps = 2, 8, 0.05
    q0 = (0.5 * seeded_normal(95, C, D)).to(DEV)
    z, u = seeded_normal(96, C, D).to(DEV), seeded_uniform(97, C).to(DEV)
    kern = build_kernel(reference_faithful=False)
    sb, ib = kern(gr.HostNoise(z, u), mm_init(q0, _funnel_logp, _monge), _funnel_logp, eps, _monge)
    s1, i1 = kern(gr.HostNoise(z[0], u[0]), mm_init(q0[0], _funnel_logp, _monge), _funnel_logp, eps, _monge)
    torch.cuda.synchronize()
    assert s1.position.shape == (D,) and s1.metric.shape == (D, D) and s1.logdensity.dim() == 0
    assert bool(i1.is_accepted) == bool(ib.is_accepted[0])
    for f in DiffusionState._fields:
        assert rel_err(getattr(s1, f), getattr(sb, f)[0]) < 1e-13, f


def test_user_functions_returning_float32_or_host_tensors_are_cast():
    """A metric written without dtype/device (float32 constants, a host
    tensor) gives the same transition as the float64 device version."""
    C, D, eps = 3, 6, 0.05
    q0 = (0.5 * seeded_normal(98, C, D)).to(DEV)
    z, u = seeded_normal(99, C, D).to(DEV), seeded_uniform(100, C).to(DEV)

    def metric64(q):
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) * (2.0 + torch.tanh(q.sum()))

    def metric32(q):
        return (torch.eye(q.shape[0], device=q.device) * (2.0 + torch.tanh(q.sum()))).float()

    def metric_host(q):  # a constant metric built on the host
        return 2.0 * torch.eye(q.shape[0], dtype=torch.float64)

    def metric_const(q):
        return 2.0 * torch.eye(q.shape[0], dtype=q.dtype, device=q.device)

    kern = build_kernel(reference_faithful=False)
    a, ia = kern(gr.HostNoise(z, u), mm_init(q0, _funnel_logp, metric64), _funnel_logp, eps, metric64)
    b, ib = kern(gr.HostNoise(z, u), mm_init(q0, _funnel_logp, metric32), _funnel_logp, eps, metric32)
    c, ic = kern(gr.HostNoise(z, u), mm_init(q0, _funnel_logp, metric_host), _funnel_logp, eps, metric_host)
    d, idd = kern(gr.HostNoise(z, u), mm_init(q0, _funnel_logp, metric_const), _funnel_logp, eps, metric_const)
    torch.cuda.synchronize()
    assert b.metric.dtype == torch.float64 and b.metric.is_cuda and c.metric.is_cuda
    assert rel_err(b.position, a.position) < 1e-5  # float32 metric values, widened
    assert rel_err(c.position, d.position) < 1e-15 and rel_err(c.Gamma, d.Gamma) < 1e-15
