"""geomjax_amd: geomjax's Riemannian integrator hot path as gfx950 HIP kernels.

Public names follow the reference package (geomjax/__init__.py:1-58) for the
samplers this build serves.
"""
from . import diagnostics, metrics, models, random, sampling
from .base import SamplingAlgorithm
from .metrics import FisherMetric, MongeMetric
from .models import LogisticBatch, LogisticRegression, NealFunnel
from .lmcmc import dynamic_lmc, lmc
from .lmcmonge import dynamic_lmc as dynamic_lmcmonge
from .lmcmonge import lmc as lmcmonge
from .rmhmc import dynamic_rmhmc, rmhmc
from .nuts import nuts as nutsrmhmc
from .lmcmonge_nuts import nutslmc as nutslmcmonge
from .lmcmc_nuts import nuts as nutslmc
from .mmala import mmala
from .glmc import glmc
from .adaptation import chees_adaptation as chees_adaptation_riemanian
from .adaptation import chees_adaptation_lmcmonge

# ``nuts`` also names the RMHMC NUTS here (the reference's ``geomjax.nuts`` is its
# Euclidean NUTS, which this build does not serve)
nuts = nutsrmhmc

__version__ = "0.1.0"

ess = diagnostics.effective_sample_size
rhat = diagnostics.potential_scale_reduction

__all__ = ["rmhmc", "dynamic_rmhmc", "nutsrmhmc", "nuts", "nutslmcmonge", "nutslmc", "mmala", "glmc", "lmc", "dynamic_lmc", "lmcmonge", "dynamic_lmcmonge",
           "chees_adaptation_riemanian", "chees_adaptation_lmcmonge", "SamplingAlgorithm", "NealFunnel", "LogisticRegression", "LogisticBatch", "MongeMetric", "FisherMetric", "metrics", "models", "random", "sampling",
           "diagnostics", "ess", "rhat", "__version__"]
