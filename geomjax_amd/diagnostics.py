"""Convergence diagnostics on device tensors: R-hat and effective sample size.

Same estimators as the reference (geomjax/diagnostics.py:25-75 and :78-209):
FFT autocovariance, Geyer's initial positive and initial monotone sequences,
tau floored at 1/log10(M N).  Written with torch so that min-ESS over
thousands of chains is computed where the samples already live (no host
round trip of the [T, C, D] sample array).  Computed in float64.
"""
import math

import torch


def _next_fast_len(n):
    """Smallest 2^a 3^b 5^c >= n (any length >= 2N gives the same autocovariance)."""
    best = 1 << (n - 1).bit_length()
    p5 = 1
    while p5 < 2 * n:
        p35 = p5
        while p35 < 2 * n:
            m = p35
            while m < n:
                m *= 2
            best = min(best, m)
            p35 *= 3
        p5 *= 5
    return best


def potential_scale_reduction(input_array, chain_axis=0, sample_axis=1):
    x = torch.as_tensor(input_array).to(torch.float64)
    if x.shape[chain_axis] < 2:
        raise ValueError("potential_scale_reduction as implemented only works for two or more chains.")
    n = x.shape[sample_axis]
    mean = x.mean(dim=sample_axis, keepdim=True)
    var = x.var(dim=sample_axis, unbiased=True, keepdim=True)
    between = n * mean.var(dim=chain_axis, unbiased=True, keepdim=True)
    within = var.mean(dim=chain_axis, keepdim=True)
    return torch.sqrt((between / within + n - 1) / n).squeeze()


def effective_sample_size(input_array, chain_axis=0, sample_axis=1):
    x = torch.as_tensor(input_array).to(torch.float64)
    sample_axis = sample_axis % x.dim()
    chain_axis = chain_axis % x.dim()
    M = x.shape[chain_axis]
    N = x.shape[sample_axis]
    if M < 2:
        raise ValueError("effective_sample_size as implemented only works for two or more chains.")
    # canonical layout [N, M, rest...]
    rest = [d for d in range(x.dim()) if d not in (sample_axis, chain_axis)]
    x = x.permute([sample_axis, chain_axis] + rest)
    mean = x.mean(dim=0, keepdim=True)
    m = _next_fast_len(2 * N)
    f = torch.fft.rfft(x - mean, n=m, dim=0)
    acov = torch.fft.irfft(f * f.conj(), n=m, dim=0)[:N] / N       # [N, M, ...]
    mean_acov = acov.mean(dim=1)                                    # [N, ...]
    var0 = mean_acov[:1] * N / (N - 1.0)
    weighted_var = var0 * (N - 1.0) / N + mean[0].var(dim=0, unbiased=True, keepdim=True)
    n_even = N - N % 2
    rho = torch.cat([torch.ones_like(var0), 1.0 - (var0 - mean_acov[1:n_even]) / weighted_var], dim=0)
    rho_even, rho_odd = rho[0::2], rho[1::2]
    # Geyer's initial positive sequence: keep pairs while every earlier pair sum is positive
    mask = torch.cumprod(((rho_even + rho_odd) > 0).to(torch.int8), dim=0).bool()
    K = rho_even.shape[0]
    max_t = (mask.to(torch.int64).sum(dim=0) - 1).clamp(min=0)     # last positive index (0 if none)
    rho_odd = torch.where(mask, rho_odd, torch.zeros_like(rho_odd))
    idx = max_t + 1
    in_range = idx < K
    gathered = torch.gather(rho_even, 0, idx.clamp(max=K - 1)[None])[0]
    mask_even = mask.clone()
    extra = torch.zeros_like(mask)
    extra.scatter_(0, idx.clamp(max=K - 1)[None], (in_range & (gathered > 0))[None])
    keep_row = torch.zeros_like(mask)
    keep_row.scatter_(0, idx.clamp(max=K - 1)[None], in_range[None])
    mask_even = torch.where(keep_row, extra, mask_even)
    rho_even = torch.where(mask_even, rho_even, torch.zeros_like(rho_even))
    # Geyer's initial monotone sequence
    s = rho_even + rho_odd
    running = torch.cummin(s, dim=0).values
    update = s > torch.cat([s[:1], running[:-1]], dim=0)
    rho_even_f = torch.where(update, running / 2.0, rho_even)
    rho_odd_f = torch.where(update, running / 2.0, rho_odd)
    tau = -1.0 + 2.0 * (rho_even_f + rho_odd_f).sum(dim=0) - torch.gather(rho_even_f, 0, idx.clamp(max=K - 1)[None])[0]
    tau = torch.clamp(tau, min=1.0 / math.log10(M * N))
    return ((M * N) / tau).squeeze()


ess = effective_sample_size
rhat = potential_scale_reduction
