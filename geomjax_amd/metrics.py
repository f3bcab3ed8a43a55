"""Metric descriptors served by the HIP kernels.

``MongeMetric`` is the metric_fn G(q) = I + alpha2 * grad l(q) grad l(q)^T
that the RMHMC configurations pass to ``rmhmc`` (the reference leaves
``metric_fn`` user-supplied, geomjax/rmhmc/rmhmc.py:286-311).  Its algebra
(Sherman-Morrison inverse, determinant lemma, rank-one Cholesky, dT/dq) is
written in closed form inside the kernels.
"""


class MongeMetric:
    kind = "monge"

    def __init__(self, alpha2=1.0):
        if not alpha2 >= 0:
            raise ValueError(f"alpha2 must be >= 0 (got {alpha2})")
        self.alpha2 = float(alpha2)

    def __repr__(self):
        return f"MongeMetric(alpha2={self.alpha2})"


class FisherMetric:
    """The Fisher information metric of the model plus its prior precision.

    For LogisticRegression: G(q) = X^T diag(s(1-s)) X + I / prior_var.
    """

    kind = "fisher"

    def __repr__(self):
        return "FisherMetric()"


def _check_alpha2(alpha2):
    if not float(alpha2) >= 0:
        raise ValueError(f"alpha2 must be >= 0 (got {alpha2})")
