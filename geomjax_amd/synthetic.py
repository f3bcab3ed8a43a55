"""Seeded synthetic problems of the logistic-regression workloads.

make_logistic: X [N, D] iid N(0, 1/D), beta* ~ N(0, I), y ~ Bernoulli(sigmoid(X beta*)),
prior N(0, 100 I) (the Bayesian logistic regression configuration).
make_uneven_batch: P posteriors generated the same way with D uniform in
{8, 16, 32, 64, 128} and N = round(exp(Uniform[ln 256, ln 4096])).
Generated on the host with numpy's PCG64 (deterministic for a seed).
"""
import math

import numpy as np


def make_logistic(N=2048, D=64, seed=2):
    rng = np.random.Generator(np.random.PCG64(seed))
    X = rng.standard_normal((N, D)) / math.sqrt(D)
    beta = rng.standard_normal(D)
    p = 1.0 / (1.0 + np.exp(-(X @ beta)))
    y = (rng.random(N) < p).astype(np.float64)
    return X, y, beta


def make_uneven_batch(P=256, seed=3, dims=(8, 16, 32, 64, 128), n_range=(256, 4096)):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(P):
        D = int(rng.choice(dims))
        N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        s = int(rng.integers(0, 2**31 - 1))
        out.append(make_logistic(N, D, s))
    return out
