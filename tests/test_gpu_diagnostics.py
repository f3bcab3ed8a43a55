"""GPU: ESS and R-hat of device tensors (the diagnostics HIP kernels) equal the
oracle (float64 FFT restatement of the reference) to 1e-10.

The samples never leave the device for the diagnostics; the oracle runs on a
host copy.  Inputs: AR(1) series (rho 0, 0.7, 0.98) in both axis orders,
float32 samples, a strided view of a chain range, an odd series length, and a
[T, C, D] = [200, 4096, 64] stack of the size the sampling benchmarks produce.
"""
import numpy as np
import pytest
import torch

from geomjax_amd import diagnostics as gd
from oracle import diagnostics as od

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _ar1(shape, sample_axis, rho, seed):
    x = np.random.default_rng(seed).standard_normal(shape)
    if rho > 0:
        x = np.moveaxis(x, sample_axis, 0)
        for t in range(1, x.shape[0]):
            x[t] = rho * x[t - 1] + np.sqrt(1 - rho * rho) * x[t]
        x = np.moveaxis(x, 0, sample_axis)
    return np.ascontiguousarray(x)


@pytest.mark.parametrize("shape,chain_axis,sample_axis", [((500, 6, 3), 1, 0), ((4, 301, 2), 0, 1)])
@pytest.mark.parametrize("rho", [0.0, 0.7, 0.98])
def test_device_ess_and_rhat_match_oracle(shape, chain_axis, sample_axis, rho):
    x = _ar1(shape, sample_axis, rho, 7)
    xd = torch.from_numpy(x).to(DEV)
    e = gd.effective_sample_size(xd, chain_axis=chain_axis, sample_axis=sample_axis)
    r = gd.potential_scale_reduction(xd, chain_axis=chain_axis, sample_axis=sample_axis)
    torch.cuda.synchronize()
    assert e.is_cuda and r.is_cuda
    e_o = od.effective_sample_size(x, chain_axis=chain_axis, sample_axis=sample_axis)
    r_o = od.potential_scale_reduction(x, chain_axis=chain_axis, sample_axis=sample_axis)
    assert np.allclose(e.cpu().numpy(), e_o, rtol=1e-10, atol=0)
    assert np.allclose(r.cpu().numpy(), r_o, rtol=1e-12, atol=0)


def test_device_ess_at_benchmark_size_matches_oracle():
    T, C, D = 200, 4096, 64
    print(f"case [T, C, D] = [{T}, {C}, {D}]", flush=True)
    x = _ar1((T, C, D), 0, 0.6, 8)
    xd = torch.from_numpy(x).to(DEV)
    e = gd.effective_sample_size(xd, chain_axis=1, sample_axis=0)
    torch.cuda.synchronize()
    e_o = od.effective_sample_size(x, chain_axis=1, sample_axis=0)
    err = float(np.max(np.abs(e.cpu().numpy() / e_o - 1)))
    print(f"  max relative difference {err:.2e}; min ESS {float(e.min()):.1f} "
          f"(AR(1) closed form {T * C * 0.4 / 1.6:.1f})", flush=True)
    assert err < 1e-10


def test_device_float32_strided_views_and_odd_lengths():
    x = _ar1((301, 40, 6), 0, 0.8, 9)
    xd = torch.from_numpy(x).to(DEV)
    # a chain range and a coordinate range of a [T, C, D] array: a non-contiguous view
    view = xd[:, 8:29, 1:5]
    e = gd.effective_sample_size(view, chain_axis=1, sample_axis=0)
    r = gd.potential_scale_reduction(view, chain_axis=1, sample_axis=0)
    torch.cuda.synchronize()
    xv = np.ascontiguousarray(x[:, 8:29, 1:5])
    assert np.allclose(e.cpu().numpy(), od.effective_sample_size(xv, chain_axis=1, sample_axis=0), rtol=1e-10)
    assert np.allclose(r.cpu().numpy(), od.potential_scale_reduction(xv, chain_axis=1, sample_axis=0), rtol=1e-12)
    # float32 samples: the estimators run in float64 on the float32 values
    x32 = x.astype(np.float32)
    e32 = gd.effective_sample_size(torch.from_numpy(x32).to(DEV), chain_axis=1, sample_axis=0)
    torch.cuda.synchronize()
    ref = od.effective_sample_size(x32.astype(np.float64), chain_axis=1, sample_axis=0)
    assert np.allclose(e32.cpu().numpy(), ref, rtol=1e-10)
