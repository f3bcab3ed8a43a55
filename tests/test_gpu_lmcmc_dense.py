"""GPU parity: LMC with a general dense user metric (lmcmc with any metric_fn).

The user functions are written once with torch operations and run on the CPU
for the oracle (oracle/lmcmc.py: full jacfwd Christoffel term, LU
factorisations, slogdet gradients) and on the GPU for the product path
(torch.func derivatives + the dense-metric HIP kernels).  Cases: Neal's funnel
with the Monge metric built by autodiff (its Christoffel term is not
symmetric, so the LU pivoting path is exercised), and the Fisher metric of
logistic regression written as a plain function, which must also agree with
the closed-form logistic LMC kernel.  Bars: 1e-10 norm-wise relative on
positions, log densities, gradients, velocities, energies and acceptance;
accept decisions identical."""
import math

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import random as gr
from geomjax_amd.lmcmc import LMCState
from geomjax_amd.lmcmc import build_kernel as lmc_kernel
from geomjax_amd.lmcmc import init as lmc_init
from geomjax_amd.synthetic import make_logistic
from oracle import lmcmc as olc
from tests.helpers import rel_err, seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def funnel_logp(q, sigma=3.0):
    x, v = q[:-1], q[-1]
    lv = -0.5 * v * v / sigma ** 2 - 0.5 * math.log(2 * math.pi * sigma ** 2)
    lx = torch.sum(-0.5 * x * x * torch.exp(-v) - 0.5 * v - 0.5 * math.log(2 * math.pi))
    return lv + lx


def monge_metric(alpha2):
    g = torch.func.grad(funnel_logp)

    def metric_fn(q):
        gq = g(q)
        return torch.eye(q.shape[0], dtype=q.dtype, device=q.device) + alpha2 * torch.outer(gq, gq)

    return metric_fn


@pytest.mark.parametrize("D,C,eps,L,a2,T,seed", [(10, 4, 0.05, 4, 1.0, 2, 1), (20, 3, 0.1, 3, 0.5, 2, 2),
                                                 (40, 2, 0.05, 2, 1.0, 1, 3),   # 40: the capacity-64 kernels
                                                 (32, 2, 0.05, 2, 1.0, 1, 4),   # the capacity-32 layout's last D
                                                 (33, 2, 0.05, 2, 1.0, 1, 5),   # the capacity-64 layout's first D
                                                 (64, 2, 0.03, 2, 1.0, 1, 6)])  # the advertised maximum
def test_funnel_monge_autodiff_matches_oracle(D, C, eps, L, a2, T, seed):
    print(f"case dense funnel-Monge D={D} C={C} eps={eps} L={L} alpha2={a2}", flush=True)
    metric_fn = monge_metric(a2)
    q0 = 0.5 * seeded_normal(10 + seed, C, D)
    so = olc.init(q0, funnel_logp)
    s = lmc_init(q0.to(DEV), funnel_logp)
    kern, okern = lmc_kernel(), olc.build_kernel()
    for t in range(T):
        z, u = seeded_normal(20 + 10 * seed + t, C, D), seeded_uniform(30 + 10 * seed + t, C)
        s1, info = kern(gr.HostNoise(z.to(DEV), u.to(DEV)), s, funnel_logp, eps, metric_fn, L)
        so1, oi = okern((z, u), so, funnel_logp, eps, metric_fn, L)
        torch.cuda.synchronize()
        print(f"  t={t}: accepted {oi.is_accepted.tolist()}", flush=True)
        assert info.is_accepted.cpu().tolist() == oi.is_accepted.tolist()
        for f in ("position", "logdensity", "logdensity_grad"):
            assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, f
        assert (s1.volume_adjustment.cpu() - so1.volume_adjustment).abs().max() < 1e-10 * max(
            1.0, so1.volume_adjustment.abs().max().item())
        assert rel_err(info.velocity, oi.velocity) < 1e-10
        assert rel_err(info.acceptance_rate, oi.acceptance_rate) < 1e-10
        assert rel_err(info.energy, oi.energy) < 1e-10
        for f in ("position", "momentum", "velocity", "logdensity_grad"):
            assert rel_err(getattr(info.proposal.state, f), getattr(oi.proposal.state, f)) < 1e-10, f
        so = so1
        s = LMCState(*(x.to(DEV) for x in so))


def test_fisher_metric_as_a_user_function_matches_the_closed_form_kernel():
    X, y, _ = make_logistic(100, 6, 3)
    Xd = torch.as_tensor(X, dtype=torch.float64, device=DEV)
    yd = torch.as_tensor(y, dtype=torch.float64, device=DEV)
    prior = 100.0

    def logp(b):
        eta = Xd @ b
        return torch.sum(yd * eta - torch.nn.functional.softplus(eta)) - 0.5 * torch.dot(b, b) / prior \
            - 0.5 * b.shape[0] * math.log(2 * math.pi * prior)

    def fisher(b):
        s = torch.sigmoid(Xd @ b)
        return (Xd.T * (s * (1 - s))) @ Xd + torch.eye(b.shape[0], dtype=b.dtype, device=b.device) / prior

    C, D, eps, L = 8, 6, 0.1, 5
    q0 = (0.3 * seeded_normal(40, C, D)).to(DEV)
    z, u = seeded_normal(41, C, D).to(DEV), seeded_uniform(42, C).to(DEV)
    model = gj.LogisticRegression(X, y, prior)
    s_user = lmc_init(q0, logp)
    s_closed = lmc_init(q0, model)
    assert rel_err(s_user.logdensity, s_closed.logdensity) < 1e-12
    a, ia = lmc_kernel()(gr.HostNoise(z, u), s_user, logp, eps, fisher, L)
    b, ib = lmc_kernel()(gr.HostNoise(z, u), s_closed, model, eps, gj.FisherMetric(), L)
    torch.cuda.synchronize()
    assert torch.equal(ia.is_accepted, ib.is_accepted)
    for f in ("position", "logdensity", "logdensity_grad", "volume_adjustment"):
        assert rel_err(getattr(a, f), getattr(b, f)) < 1e-10, f
    assert rel_err(ia.velocity, ib.velocity) < 1e-10
    assert rel_err(ia.energy, ib.energy) < 1e-10


def test_dense_sampler_api_runs_with_device_draws():
    metric_fn = monge_metric(1.0)
    alg = gj.lmc(funnel_logp, 0.05, metric_fn, 4)
    s = alg.init(0.5 * seeded_normal(50, 32, 10).to(DEV))
    res = alg.sample_fn(gr.key(3), s, 5)
    torch.cuda.synchronize()
    assert res.samples.shape == (5, 32, 10)
    assert torch.isfinite(res.samples).all()
    assert 0.0 < float(res.acceptance_rate.mean()) <= 1.0


def test_per_chain_step_counts_match_the_oracle_chain_by_chain():
    D, C, eps = 10, 3, 0.05
    metric_fn = monge_metric(1.0)
    q0 = 0.5 * seeded_normal(60, C, D)
    z, u = seeded_normal(61, C, D), seeded_uniform(62, C)
    counts = torch.tensor([3, 0, 5])
    s1, info = lmc_kernel()(gr.HostNoise(z.to(DEV), u.to(DEV)), lmc_init(q0.to(DEV), funnel_logp), funnel_logp, eps,
                            metric_fn, counts.to(DEV), max_num_integration_steps=5)
    torch.cuda.synchronize()
    for c in range(C):
        so1, oi = olc.build_kernel()((z[c:c + 1], u[c:c + 1]), olc.init(q0[c:c + 1], funnel_logp), funnel_logp, eps,
                                     metric_fn, int(counts[c]))
        assert bool(info.is_accepted[c]) == bool(oi.is_accepted[0]), c
        assert rel_err(s1.position[c:c + 1], so1.position) < 1e-10, c
        assert abs(float(s1.volume_adjustment[c]) - float(so1.volume_adjustment[0])) < 1e-10, c
        assert rel_err(info.acceptance_rate[c:c + 1], oi.acceptance_rate) < 1e-10, c


def _skewed_metric(q):
    """G = A A^T + I/10 with A = I + 3 (subdiagonal) + diag(tanh(q))/10: column 0 of
    G holds 1.01 + ... on the diagonal and ~3 below it, so LU with partial
    pivoting of G +- (eps/2) Omega~ must swap rows."""
    D = q.shape[0]
    A = torch.eye(D, dtype=q.dtype, device=q.device) + 3.0 * torch.diag(torch.ones(D - 1, dtype=q.dtype,
                                                                                   device=q.device), -1)
    A = A + 0.1 * torch.diag(torch.tanh(q))
    return A @ A.T + 0.1 * torch.eye(D, dtype=q.dtype, device=q.device)


def _gauss_logp(q):
    return -0.5 * torch.dot(q, q)


@pytest.mark.parametrize("D", [6, 33])
def test_half_step_with_row_swaps_matches_oracle(D):
    print(f"case dense lmc skewed metric D={D} (LU row swaps)", flush=True)
    C, eps, L = 3, 0.05, 2
    q0 = 0.5 * seeded_normal(60 + D, C, D)
    _, piv = torch.linalg.lu_factor(_skewed_metric(q0[0]))
    swaps = int((piv.cpu() != torch.arange(1, D + 1, dtype=piv.dtype)).sum())
    assert swaps > 0  # the case really pivots
    so, s = olc.init(q0, _gauss_logp), lmc_init(q0.to(DEV), _gauss_logp)
    z, u = seeded_normal(61 + D, C, D), seeded_uniform(62 + D, C)
    s1, info = lmc_kernel()(gr.HostNoise(z.to(DEV), u.to(DEV)), s, _gauss_logp, eps, _skewed_metric, L)
    so1, oi = olc.build_kernel()((z, u), so, _gauss_logp, eps, _skewed_metric, L)
    torch.cuda.synchronize()
    print(f"  row swaps in LU(G(q0)): {swaps}; accepted {oi.is_accepted.tolist()}", flush=True)
    assert info.is_accepted.cpu().tolist() == oi.is_accepted.tolist()
    for f in ("position", "logdensity", "logdensity_grad"):
        assert rel_err(getattr(s1, f), getattr(so1, f)) < 1e-10, f
    assert rel_err(info.velocity, oi.velocity) < 1e-10
    assert rel_err(info.energy, oi.energy) < 1e-10
    for f in ("position", "momentum", "velocity"):
        assert rel_err(getattr(info.proposal.state, f), getattr(oi.proposal.state, f)) < 1e-10, f
    assert (info.proposal.state.volume_adjustment.cpu() - oi.proposal.state.volume_adjustment).abs().max() < 1e-10


def test_single_chain_state_equals_the_batched_row():
    """A [D] position takes the dense path as a batch of one chain; the result
    equals row 0 of the same step run batched, with unbatched shapes."""
    D, eps, L = 10, 0.05, 3
    metric_fn = monge_metric(1.0)
    q0 = (0.5 * seeded_normal(70, 2, D)).to(DEV)
    z, u = seeded_normal(71, 2, D).to(DEV), seeded_uniform(72, 2).to(DEV)
    kern = lmc_kernel()
    sb, ib = kern(gr.HostNoise(z, u), lmc_init(q0, funnel_logp), funnel_logp, eps, metric_fn, L)
    s1, i1 = kern(gr.HostNoise(z[0], u[0]), lmc_init(q0[0], funnel_logp), funnel_logp, eps, metric_fn, L)
    torch.cuda.synchronize()
    assert s1.position.shape == (D,) and s1.logdensity.dim() == 0 and i1.velocity.shape == (D,)
    assert bool(i1.is_accepted) == bool(ib.is_accepted[0])
    for f in ("position", "logdensity", "logdensity_grad", "volume_adjustment"):
        assert rel_err(getattr(s1, f), getattr(sb, f)[0]) < 1e-13, f
    assert rel_err(i1.proposal.state.position, ib.proposal.state.position[0]) < 1e-13
