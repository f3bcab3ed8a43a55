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
