"""Shared helpers of the test-suite (comparison metrics, seeded inputs)."""
import numpy as np
import torch


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error, 0 for two zero arrays)."""
    a = torch.as_tensor(a).detach().double().cpu()
    b = torch.as_tensor(b).detach().double().cpu()
    den = b.abs().max().item() if b.numel() else 0.0
    num = (a - b).abs().max().item() if b.numel() else 0.0
    if den == 0.0:
        return num
    return num / den


def seeded_normal(seed, *shape):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, generator=g, dtype=torch.float64)


def seeded_uniform(seed, *shape):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(*shape, generator=g, dtype=torch.float64)
