"""The C ABI library: loads without a GPU and exports every declared symbol (CPU)."""
import os
import re

from geomjax_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "geomjax_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gjx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gjx_rmhmc_monge_funnel_step" in names
    assert len(names) >= 4


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(_lib.exported_symbols())


def test_version_and_error_strings():
    lib = _lib.load()
    assert lib.gjx_version().decode().startswith("geomjax_amd")
    assert isinstance(lib.gjx_last_error(), bytes)


def test_invalid_arguments_are_rejected_without_a_gpu():
    """Validation happens on the host before any HIP call."""
    lib = _lib.load()
    p = _lib.FunnelParams(num_chains=0, dim=10, dtype=0, num_integration_steps=4, max_iters=8, num_transitions=1,
                          step_size=0.1, alpha2=1.0, sigma=3.0, convergence_tol=1e-6, divergence_tol=1e10,
                          divergence_threshold=1000.0)
    io = _lib.RmhmcStepIO()
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -1
    assert "num_chains" in _lib.last_error()
    p.num_chains, p.dim = 4, 1
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -1
    p.dim, p.lanes_per_chain = 10, 3
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -1
    assert "lanes_per_chain" in _lib.last_error()
    p.lanes_per_chain, p.dim = 4, 1000
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -2
    p.dim, p.alpha2 = 10, -1.0
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -1
    p.alpha2 = 1.0
    assert lib.gjx_rmhmc_monge_funnel_step(p, io, None) == -1  # buffers missing
    assert "position" in _lib.last_error()


def _pack_size_numpy(N, D):
    """Bytes of the packed float32 table: header of 40-byte entries rounded up to 256 bytes, then
    per posterior three fragment layouts of npad x 32nb words (Gram, rows, rows split into bf16
    pairs) and npad (rounded up to 64) labels."""
    P = len(N)
    off = -(-P * 40 // 256) * 64
    for n, d in zip(N, D):
        npad, nb = -(-n // 64) * 64, 1 if d <= 32 else (2 if d <= 64 else 4)
        off += 3 * npad * 32 * nb + -(-npad // 64) * 64
    return off * 4


def test_logistic_pack_and_workspace_sizes():
    import ctypes

    lib = _lib.load()
    for N, D in [([2048], [64]), ([200, 4096, 333], [8, 64, 33]), ([31], [1]), ([500, 77], [128, 65])]:
        n = (ctypes.c_int32 * len(N))(*N)
        d = (ctypes.c_int32 * len(D))(*D)
        out = ctypes.c_uint64()
        assert lib.gjx_logistic_pack_size(len(N), n, d, ctypes.byref(out)) == 0
        assert out.value == _pack_size_numpy(N, D)
    n, d = (ctypes.c_int32 * 1)(100), (ctypes.c_int32 * 1)(129)
    assert lib.gjx_logistic_pack_size(1, n, d, ctypes.byref(out)) == -2
    p = _lib.LogisticParams(num_chains=7, dim=64, dtype=_lib.GJX_F32)
    data = _lib.LogisticData(num_posteriors=1, max_rows=2000, max_features=64)
    assert lib.gjx_logistic_workspace_size(p, data, ctypes.byref(out)) == 0
    assert out.value == 7 * 5 * 2048 * 4


def test_logistic_float32_requires_packed_table_and_workspace():
    lib = _lib.load()
    p = _lib.LogisticParams(num_chains=4, dim=8, dtype=_lib.GJX_F32, num_integration_steps=2, max_iters=2,
                            num_transitions=1, step_size=0.1, prior_var=100.0)
    data = _lib.LogisticData(num_posteriors=1, max_rows=16, max_features=8, num_rows=1, num_features=1,
                             x_offset=1, y_offset=1, x=1, y=1)
    io = _lib.RmhmcStepIO()
    assert lib.gjx_rmhmc_fisher_logistic_step(p, data, io, None) == -1
    assert "packed" in _lib.last_error()
    data.packed, data.workspace, data.workspace_bytes = 1, 1, 10
    assert lib.gjx_rmhmc_fisher_logistic_step(p, data, io, None) == -1
    assert "workspace" in _lib.last_error()
    # the tuned LMC entry points check the same (their io is checked after)
    data.workspace_bytes = 10
    assert lib.gjx_lmc_fisher_logistic_step(p, data, _lib.LmcmcStepIO(), None) == -1
    assert "workspace" in _lib.last_error()
    assert lib.gjx_lmc_fisher_logistic_sample(p, data, _lib.LmcmcSampleIO(), None) == -1
    assert "workspace" in _lib.last_error()
    # entry points without a float32 path say so
    assert lib.gjx_mmala_fisher_logistic_step(p, data, 1, _lib.MmalaStepIO(), None) == -2
    assert "float32" in _lib.last_error()


def test_logistic_feature_capacity_per_path():
    """RMHMC (float64 and float32) takes up to 128 features, LMC (both) up to 64,
    the tangent 64 in float64 and 128 in float32."""
    lib = _lib.load()

    def step(fn, io, dtype, D):
        p = _lib.LogisticParams(num_chains=4, dim=D, dtype=dtype, num_integration_steps=2, max_iters=2,
                                num_transitions=1, step_size=0.1, prior_var=100.0)
        data = _lib.LogisticData(num_posteriors=1, max_rows=16, max_features=D)  # data arrays missing
        return fn(p, data, io, None)

    rm, lmc = lib.gjx_rmhmc_fisher_logistic_step, lib.gjx_lmc_fisher_logistic_step
    rio, lio = _lib.RmhmcStepIO(), _lib.LmcmcStepIO()
    assert step(rm, rio, _lib.GJX_F64, 128) == -1 and "data arrays" in _lib.last_error()
    assert step(rm, rio, _lib.GJX_F64, 129) == -2
    assert step(rm, rio, _lib.GJX_F32, 128) == -1
    assert step(rm, rio, _lib.GJX_F32, 129) == -2
    assert step(lmc, lio, _lib.GJX_F64, 64) == -1
    assert step(lmc, lio, _lib.GJX_F64, 65) == -2
    assert step(lmc, lio, _lib.GJX_F32, 64) == -1
    assert step(lmc, lio, _lib.GJX_F32, 65) == -2
    # the tuning tangent: float64 up to 64 features, float32 up to 128
    tan, tio = lib.gjx_rmhmc_fisher_logistic_tangent, _lib.TangentIO()
    assert step(tan, tio, _lib.GJX_F64, 64) == -1
    assert step(tan, tio, _lib.GJX_F64, 65) == -2
    assert step(tan, tio, _lib.GJX_F32, 128) == -1
    assert step(tan, tio, _lib.GJX_F32, 129) == -2


def test_mmala_and_logistic_nuts_entry_points_validate_on_the_host():
    """MMALA and the logistic NUTS entry points: MMALA float64 only (both NUTS
    entry points also float32, with the tuned path's packed table), D <= 64,
    a positive step size for MMALA, required buffers, z and u given together."""
    lib = _lib.load()

    def params(D, dtype=_lib.GJX_F64, step=0.1):
        return _lib.LogisticParams(num_chains=4, dim=D, dtype=dtype, num_integration_steps=0, max_iters=0,
                                   num_transitions=3, step_size=step, prior_var=100.0)

    def data(D):
        return _lib.LogisticData(num_posteriors=1, max_rows=16, max_features=D, num_rows=1, num_features=1,
                                 x_offset=1, y_offset=1, x=1, y=1)

    mstep, msample = lib.gjx_mmala_fisher_logistic_step, lib.gjx_mmala_fisher_logistic_sample
    assert mstep(params(8, _lib.GJX_F32), data(8), 1, _lib.MmalaStepIO(), None) == -2
    assert "float32" in _lib.last_error()
    assert mstep(params(65), data(65), 1, _lib.MmalaStepIO(), None) == -2
    assert mstep(params(8, step=0.0), data(8), 1, _lib.MmalaStepIO(), None) == -1
    assert "step_size" in _lib.last_error()
    assert mstep(params(8), data(8), 1, _lib.MmalaStepIO(), None) == -1
    assert "position" in _lib.last_error()
    io = _lib.MmalaStepIO(position=1, logdensity=1, new_position=1, new_logdensity=1, z=1)
    assert mstep(params(8), data(8), 0, io, None) == -1
    assert "z and u" in _lib.last_error()
    sio = _lib.MmalaSampleIO(position=1, logdensity=1)
    assert msample(params(8), data(8), 1, sio, None) == -1
    assert "final_position" in _lib.last_error()
    assert lib.gjx_mmala_fisher_logistic_init(params(65), data(65), 1, 1, 1, None, None, None) == -2
    assert lib.gjx_mmala_fisher_logistic_init(params(8), data(8), 1, None, 1, None, None, None) == -1
    for fn, io_t in ((lib.gjx_rmhmc_fisher_logistic_nuts, _lib.NutsIO), (lib.gjx_lmc_fisher_logistic_nuts,
                                                                        _lib.LmcNutsIO)):
        assert fn(params(65), data(65), 5, _lib.GJX_UTURN["euc"], io_t(), None) == -2
        assert fn(params(8), data(8), 5, 7, io_t(), None) == -1  # unknown criterion
    nuts_rm = lib.gjx_rmhmc_fisher_logistic_nuts
    assert nuts_rm(params(8, _lib.GJX_F32), data(8), 5, _lib.GJX_UTURN["euc"], _lib.NutsIO(), None) == -1
    assert "packed" in _lib.last_error()
    d32 = data(8)
    d32.packed, d32.workspace, d32.workspace_bytes = 1, 1, 1 << 40
    assert nuts_rm(params(8, _lib.GJX_F32), d32, 5, _lib.GJX_UTURN["euc"], _lib.NutsIO(), None) == -1
    assert "position" in _lib.last_error()
    assert lib.gjx_lmc_fisher_logistic_nuts(params(8, _lib.GJX_F32), d32, 5, _lib.GJX_UTURN["euc"], _lib.LmcNutsIO(),
                                            None) == -1
    assert "position" in _lib.last_error()


def test_dense_metric_entry_points_validate_on_the_host():
    lib = _lib.load()
    io = _lib.DenseMetricIO(num_chains=0, dim=8)
    assert lib.gjx_dense_metric_position(io, None) == -1
    assert "num_chains" in _lib.last_error()
    io.num_chains, io.dim = 4, 65
    assert lib.gjx_dense_metric_position(io, None) == -2
    io.dim = 8
    assert lib.gjx_dense_metric_position(io, None) == -1
    assert "metric" in _lib.last_error()
    io.metric, io.metric_jacobian, io.velocity_out = 1, 1, 1
    assert lib.gjx_dense_metric_position(io, None) == -1
    assert "needs z" in _lib.last_error()
    io.velocity_out = None
    assert lib.gjx_dense_lmc_half_step(io, None) == -1
    assert "required" in _lib.last_error()


def test_dense_mmala_drift_validates_on_the_host():
    lib = _lib.load()
    io = _lib.DenseMetricIO(num_chains=2, dim=4, metric=1, gamma=1)
    assert lib.gjx_dense_mmala_drift(io, None) == -1
    assert "metric_jacobian" in _lib.last_error()
    io.gamma, io.drift = None, 1
    assert lib.gjx_dense_mmala_drift(io, None) == -1
    assert "logdensity_grad" in _lib.last_error()
    io.drift, io.chol_inv_z = None, 1
    assert lib.gjx_dense_mmala_drift(io, None) == -1
    assert "needs z" in _lib.last_error()


def test_dense_rmhmc_field_validates_on_the_host():
    lib = _lib.load()
    io = _lib.DenseMetricIO(num_chains=2, dim=4, metric=1)
    assert lib.gjx_dense_rmhmc_field(io, None) == -1
    assert "momentum or z" in _lib.last_error()
    io.z, io.dtdq = 1, 1
    assert lib.gjx_dense_rmhmc_field(io, None) == -1
    assert "metric_jacobian" in _lib.last_error()


def test_waves_per_chain_field_is_validated_on_the_host():
    """Each 4-bit field of waves_per_chain must be 0, 1 or 2; checked before any launch."""
    lib = _lib.load()
    data = _lib.LogisticData(num_posteriors=1, max_rows=16, max_features=8, num_rows=1, num_features=1,
                             x_offset=1, y_offset=1, x=1, y=1)
    for bad in (0x3, 0x30, 0x100, -1):
        p = _lib.LogisticParams(num_chains=4, dim=8, dtype=_lib.GJX_F32, num_integration_steps=2, max_iters=2,
                                num_transitions=1, step_size=0.1, prior_var=100.0, waves_per_chain=bad)
        assert lib.gjx_rmhmc_fisher_logistic_step(p, data, _lib.RmhmcStepIO(), None) == _lib.GJX_EINVAL
        assert "waves_per_chain" in _lib.last_error()
    p = _lib.LogisticParams(num_chains=4, dim=8, dtype=_lib.GJX_F32, num_integration_steps=2, max_iters=2,
                            num_transitions=1, step_size=0.1, prior_var=100.0, waves_per_chain=0x21)
    assert lib.gjx_rmhmc_fisher_logistic_step(p, data, _lib.RmhmcStepIO(), None) == _lib.GJX_EINVAL
    assert "packed" in _lib.last_error()  # the field passed; the next check fails


def test_waves_argument_encoding():
    import pytest
    import torch

    from geomjax_amd.rmhmc import _waves

    assert _waves(3, None, 4, torch.float64) == 0  # float64 calls ignore it
    assert _waves(1, None, 4, torch.float32) == 0x11
    assert _waves(2, None, 4, torch.float32) == 0x22
    assert _waves((2, 1), None, 4, torch.float32) == 0x12
    with pytest.raises(ValueError):
        _waves(3, None, 4, torch.float32)
    with pytest.raises(ValueError):
        _waves((1, 4), None, 4, torch.float32)


def test_diagnostics_entry_points_validate_on_the_host():
    import ctypes

    lib = _lib.load()
    nbytes = ctypes.c_uint64(0)
    a = _lib.DiagnosticsArgs(num_samples=100, num_chains=8, num_coords=3, dtype=_lib.GJX_F32, stride_sample=24,
                             stride_chain=3, stride_coord=1)
    assert lib.gjx_diagnostics_workspace_size(ctypes.byref(a), ctypes.byref(nbytes)) == _lib.GJX_OK
    assert nbytes.value >= 8 * 3 * 100  # at least the per-coordinate lag sums
    bad = _lib.DiagnosticsArgs(num_samples=100, num_chains=1, num_coords=3, dtype=_lib.GJX_F32)
    assert lib.gjx_diagnostics_workspace_size(ctypes.byref(bad), ctypes.byref(nbytes)) == _lib.GJX_EINVAL
    assert "num_chains" in _lib.last_error()
    long = _lib.DiagnosticsArgs(num_samples=10 ** 6, num_chains=4, num_coords=1, dtype=_lib.GJX_F64)
    assert lib.gjx_diagnostics_workspace_size(ctypes.byref(long), ctypes.byref(nbytes)) == _lib.GJX_EUNSUPPORTED
    # no workspace: refused before any launch
    assert lib.gjx_effective_sample_size(ctypes.byref(a), 1, None, 0, 1, None, None) == _lib.GJX_EINVAL
    assert "workspace" in _lib.last_error()
