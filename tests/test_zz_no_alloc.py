"""GPU: the allocation-free tuned RMHMC entry point
(gjx_rmhmc_fisher_logistic_step_noalloc, scratch in a static per-GPU array)
equals gjx_rmhmc_fisher_logistic_step (workspace from the caller) bit for bit
on the same inputs, on two calls in a row -- the second with fewer chains and
rows, so it runs on scratch the first call left behind -- and refuses a call
whose scratch would not fit, before any launch.
"""
import ctypes

import pytest
import torch

import geomjax_amd as gj
from geomjax_amd import _lib
from geomjax_amd.models import logistic_params
from geomjax_amd.synthetic import make_logistic
from tests.helpers import seeded_normal, seeded_uniform

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
OUT_FIELDS = ("new_position", "new_logdensity", "new_logdensity_grad", "momentum", "acceptance_rate",
              "is_accepted", "is_divergent", "energy", "proposal_position", "proposal_momentum",
              "proposal_velocity", "proposal_logdensity", "proposal_logdensity_grad", "proposal_weight",
              "proposal_sum_log_p_accept", "fp_iters", "fp_norm", "fp_success")


def _outputs(C, D, L):
    vec = lambda: torch.full((C, D), float("nan"), dtype=torch.float32, device=DEV)
    sca = lambda dt=torch.float64: torch.full((C,), 7, dtype=dt, device=DEV)
    return dict(new_position=vec(), new_logdensity=sca(), new_logdensity_grad=vec(), momentum=vec(),
                acceptance_rate=sca(), is_accepted=sca(torch.uint8), is_divergent=sca(torch.uint8), energy=sca(),
                proposal_position=vec(), proposal_momentum=vec(), proposal_velocity=vec(),
                proposal_logdensity=sca(), proposal_logdensity_grad=vec(), proposal_weight=sca(),
                proposal_sum_log_p_accept=sca(),
                fp_iters=torch.full((C, L), -1, dtype=torch.int32, device=DEV),
                fp_norm=torch.full((C, L), -1.0, dtype=torch.float64, device=DEV),
                fp_success=torch.full((C, L), 9, dtype=torch.uint8, device=DEV))


def _call(entry, model, q, logp, grad, z, u, L, eps, workspace=True):
    C, D = q.shape
    params = logistic_params(model, C, L=L, step_size=eps, max_iters=6, convergence_tol=1e-6, divergence_tol=1e10,
                             divergence_threshold=1000.0, dtype=torch.float32)
    data = model.data(dtype=torch.float32, num_chains=C)
    if not workspace:
        data.workspace, data.workspace_bytes = None, 0
    out = _outputs(C, D, L)
    io = _lib.RmhmcStepIO(position=q.data_ptr(), logdensity=logp.data_ptr(), logdensity_grad=grad.data_ptr(),
                          z=z.data_ptr(), u=u.data_ptr(), num_integration_steps=None,
                          **{k: v.data_ptr() for k, v in out.items()})
    rc = getattr(_lib.load(), entry)(ctypes.byref(params), ctypes.byref(data), ctypes.byref(io),
                                     _lib.stream_handle(q.device))
    return rc, out


def _inputs(N, D, C, seed):
    X, y, _ = make_logistic(N, D, seed)
    model = gj.LogisticRegression(X, y, 100.0)
    s = gj.rmhmc.init(seeded_normal(seed, C, D).float().to(DEV), model)
    z = seeded_normal(100 + seed, C, D).float().to(DEV)
    u = seeded_uniform(200 + seed, C).to(DEV)
    return model, s.position.contiguous(), s.logdensity.contiguous(), s.logdensity_grad.contiguous(), z, u


@pytest.mark.parametrize("first,second", [((2048, 64, 512), (700, 40, 96)), ((1024, 128, 64), (300, 20, 40))])
def test_noalloc_step_equals_the_workspace_step_on_two_calls(first, second):
    L, eps = 6, 0.1
    for k, (N, D, C) in enumerate((first, second)):
        print(f"case noalloc call {k}: N={N} D={D} C={C}", flush=True)
        model, q, logp, grad, z, u = _inputs(N, D, C, 40 + k)
        rc_ref, ref = _call("gjx_rmhmc_fisher_logistic_step", model, q, logp, grad, z, u, L, eps)
        rc, out = _call("gjx_rmhmc_fisher_logistic_step_noalloc", model, q, logp, grad, z, u, L, eps,
                        workspace=False)
        torch.cuda.synchronize()
        assert rc_ref == 0 and rc == 0, _lib.last_error()
        for name in OUT_FIELDS:
            a, b = out[name], ref[name]
            assert torch.equal(a, b) or (a.is_floating_point() and torch.equal(torch.isnan(a), torch.isnan(b))
                                         and torch.equal(a[~torch.isnan(a)], b[~torch.isnan(b)])), name
        assert torch.isfinite(out["new_position"]).all()
        print(f"  accepted {int(out['is_accepted'].sum())}/{C}, identical to the workspace entry point", flush=True)


def test_noalloc_step_refuses_scratch_beyond_its_size():
    N, D, C = 4096, 16, 1024  # 1024 x 5 x 4096 x 4 bytes = 80 MiB > 64 MiB
    model, q, logp, grad, z, u = _inputs(N, D, C, 50)
    rc, out = _call("gjx_rmhmc_fisher_logistic_step_noalloc", model, q, logp, grad, z, u, 4, 0.1, workspace=False)
    torch.cuda.synchronize()
    assert rc == _lib.GJX_EUNSUPPORTED
    assert "scratch" in _lib.last_error()
    assert torch.isnan(out["new_position"]).all()  # nothing was launched
