"""geomjax_amd: geomjax's Riemannian integrator hot path as gfx950 HIP kernels.

Public names follow the reference package (geomjax/__init__.py:1-58) for the
samplers this build serves.
"""
from . import diagnostics, metrics, models, random, sampling
from .base import SamplingAlgorithm
from .metrics import FisherMetric, MongeMetric
from .models import LogisticBatch, LogisticRegression, NealFunnel
from .lmcmc import lmc
from .lmcmonge import lmc as lmcmonge
from .rmhmc import dynamic_rmhmc, rmhmc
from .nuts import nuts

__version__ = "0.1.0"

ess = diagnostics.effective_sample_size
rhat = diagnostics.potential_scale_reduction

__all__ = ["rmhmc", "dynamic_rmhmc", "nuts", "lmc", "lmcmonge", "SamplingAlgorithm", "NealFunnel", "LogisticRegression", "LogisticBatch", "MongeMetric", "FisherMetric", "metrics", "models", "random", "sampling",
           "diagnostics", "ess", "rhat", "__version__"]
