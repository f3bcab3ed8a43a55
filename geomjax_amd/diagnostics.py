"""Convergence diagnostics: R-hat and effective sample size.

Same estimators as the reference (geomjax/diagnostics.py:25-75 and :78-209):
autocovariances of every chain, Geyer's initial positive and initial
monotone sequences, tau floored at 1/log10(M N).  Computed in float64.

Samples on a HIP device are reduced there by the library's diagnostics
kernels (``gjx_effective_sample_size``: lag sums of each chain tile in LDS,
then the Geyer steps per coordinate), so min-ESS over thousands of chains
needs no host round trip of the [T, C, D] sample array.  Host tensors and
arrays take the same estimators written with torch (FFT autocovariance).
"""
import ctypes
import math

import torch

from . import _lib


def _device_diagnostics(x, chain_axis, sample_axis, want_ess, want_rhat):
    """ESS and/or R-hat of a CUDA tensor through the HIP kernels; results shaped
    like the reference's (the chain and sample axes removed, squeezed)."""
    sample_axis = sample_axis % x.dim()
    chain_axis = chain_axis % x.dim()
    rest = [d for d in range(x.dim()) if d not in (sample_axis, chain_axis)]
    rest_shape = [x.shape[d] for d in rest]
    v = x.permute([sample_axis, chain_axis] + rest)
    if v.dtype not in (torch.float32, torch.float64):
        v = v.to(torch.float64)
    N, M = v.shape[0], v.shape[1]
    v = v.reshape(N, M, -1)  # a view where the layout allows, else a device copy
    D = v.shape[2]
    args = _lib.DiagnosticsArgs(num_samples=N, num_chains=M, num_coords=D,
                                dtype=_lib.GJX_F32 if v.dtype == torch.float32 else _lib.GJX_F64,
                                stride_sample=v.stride(0), stride_chain=v.stride(1), stride_coord=v.stride(2))
    lib = _lib.load()
    nbytes = ctypes.c_uint64(0)
    _lib.check(lib.gjx_diagnostics_workspace_size(ctypes.byref(args), ctypes.byref(nbytes)),
               "gjx_diagnostics_workspace_size")
    ws = torch.empty(max(int(nbytes.value), 8), dtype=torch.uint8, device=x.device)
    ess = torch.empty(D, dtype=torch.float64, device=x.device) if want_ess else None
    rhat = torch.empty(D, dtype=torch.float64, device=x.device) if want_rhat else None
    rc = lib.gjx_effective_sample_size(ctypes.byref(args), v.data_ptr(), ws.data_ptr(), ws.numel(), _lib.ptr(ess),
                                       _lib.ptr(rhat), _lib.stream_handle(x.device))
    _lib.check(rc, "gjx_effective_sample_size")
    shape = lambda t: None if t is None else t.reshape(rest_shape).squeeze()
    return shape(ess), shape(rhat)


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
    x = torch.as_tensor(input_array)
    if x.shape[chain_axis] < 2:
        raise ValueError("potential_scale_reduction as implemented only works for two or more chains.")
    if x.is_cuda:
        return _device_diagnostics(x, chain_axis, sample_axis, False, True)[1]
    x = x.to(torch.float64)
    n = x.shape[sample_axis]
    mean = x.mean(dim=sample_axis, keepdim=True)
    var = x.var(dim=sample_axis, unbiased=True, keepdim=True)
    between = n * mean.var(dim=chain_axis, unbiased=True, keepdim=True)
    within = var.mean(dim=chain_axis, keepdim=True)
    return torch.sqrt((between / within + n - 1) / n).squeeze()


def effective_sample_size(input_array, chain_axis=0, sample_axis=1):
    x = torch.as_tensor(input_array)
    sample_axis = sample_axis % x.dim()
    chain_axis = chain_axis % x.dim()
    M = x.shape[chain_axis]
    N = x.shape[sample_axis]
    if M < 2:
        raise ValueError("effective_sample_size as implemented only works for two or more chains.")
    if x.is_cuda:
        return _device_diagnostics(x, chain_axis, sample_axis, True, False)[0]
    x = x.to(torch.float64)
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
