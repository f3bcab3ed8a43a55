"""Host-side logic of the user-metric path (CPU): which (log density, metric)
pairs take it, the Monge metric it builds from a user log density, and the
embedding of a diagonal metric_fn as diag(m) with its Jacobian."""
import torch

import geomjax_amd as gj
from geomjax_amd.lmcmc_dense import _metric, is_dense_metric_fn, resolve_metric_fn
from tests.helpers import rel_err, seeded_normal


def _gauss(q):
    return -0.5 * torch.dot(q, q)


def test_routing():
    eye = lambda q: torch.eye(q.shape[0], dtype=q.dtype)
    assert is_dense_metric_fn(_gauss, eye)
    assert is_dense_metric_fn(_gauss, gj.MongeMetric(1.0))
    assert not is_dense_metric_fn(_gauss, gj.FisherMetric())
    assert not is_dense_metric_fn(gj.NealFunnel(5), eye)
    assert not is_dense_metric_fn(gj.NealFunnel(5), gj.MongeMetric(1.0))


def test_monge_metric_from_a_user_logdensity():
    """grad of -|q|^2/2 is -q, so the Monge metric is I + alpha2 q q^T."""
    q = seeded_normal(1, 5)
    G = resolve_metric_fn(_gauss, gj.MongeMetric(0.7))(q)
    assert rel_err(G, torch.eye(5, dtype=torch.float64) + 0.7 * torch.outer(q, q)) < 1e-15
    f = lambda x: x
    assert resolve_metric_fn(_gauss, f) is f


def test_diagonal_metric_embedding_equals_the_dense_jacobian_of_diag():
    def m(q):
        return 1.0 + torch.sin(q) ** 2 + q[0] * q

    q = seeded_normal(2, 3, 5)
    G, dG = _metric(m, q)
    G2, dG2 = _metric(lambda x: torch.diag(m(x)), q)
    assert torch.equal(G, G2) and torch.equal(dG, dG2)
    assert G.shape == (3, 5, 5) and dG.shape == (3, 5, 5, 5)


def test_user_function_outputs_are_widened_to_float64():
    """Common torch idioms give float32 (torch.ones / torch.eye without dtype);
    the kernels read float64, so the outputs are widened before any launch."""
    from geomjax_amd.lmcmc_dense import _value_and_grad

    q = seeded_normal(3, 4, 5)
    G, dG = _metric(lambda x: torch.eye(x.shape[0]) * (1.0 + x.sum()), q)
    assert G.dtype == torch.float64 and dG.dtype == torch.float64
    assert G.shape == (4, 5, 5) and dG.shape == (4, 5, 5, 5)
    ref = torch.eye(5, dtype=torch.float64) * (1.0 + q.sum(1))[:, None, None]
    assert rel_err(G, ref) < 1e-6
    Gd, dGd = _metric(lambda x: torch.ones(x.shape[0]) + x * x, q)  # a diagonal float32 metric
    assert Gd.dtype == torch.float64 and dGd.shape == (4, 5, 5, 5)
    l, g = _value_and_grad(lambda x: (x.float() ** 2).sum(), q)
    assert l.dtype == torch.float64 and g.dtype == torch.float64 and l.shape == (4,) and g.shape == (4, 5)


def test_user_function_outputs_of_the_wrong_shape_or_type_raise():
    import pytest

    from geomjax_amd.lmcmc_dense import _value_and_grad

    q = seeded_normal(4, 3, 5)
    with pytest.raises(ValueError):
        _metric(lambda x: torch.eye(x.shape[0] + 1, dtype=x.dtype), q)  # [D+1, D+1]
    with pytest.raises(ValueError):
        _metric(lambda x: torch.ones(2, 2, 2, dtype=x.dtype) * x.sum(), q)  # not a matrix
    # integer metrics and non-scalar log densities are refused (by this check or by torch.func)
    with pytest.raises((TypeError, RuntimeError)):
        _metric(lambda x: (torch.eye(x.shape[0], dtype=x.dtype) * x.sum()).to(torch.int64), q)
    with pytest.raises((ValueError, RuntimeError)):
        _value_and_grad(lambda x: x * 2.0, q)


def test_ptr_refuses_host_tensors():
    import pytest

    from geomjax_amd import _lib

    assert _lib.ptr(None) is None
    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(3, dtype=torch.float64))


def test_single_chain_batching_round_trip():
    from geomjax_amd.lmcmc import LMCState
    from geomjax_amd.lmcmc_dense import as_batch, unbatch

    s = LMCState(seeded_normal(5, 4), torch.tensor(1.5, dtype=torch.float64), seeded_normal(6, 4),
                 torch.tensor(0.0, dtype=torch.float64))
    b, single = as_batch(s)
    assert single and b.position.shape == (1, 4) and b.logdensity.shape == (1,)
    back = unbatch((b, [b.position], {"x": b.logdensity}))
    assert torch.equal(back[0].position, s.position) and back[0].logdensity.dim() == 0
    assert back[1][0].shape == (4,) and back[2]["x"].dim() == 0
    same, single = as_batch(b)
    assert not single and same is b
