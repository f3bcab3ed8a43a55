"""ctypes binding of the C ABI in include/geomjax_amd.h.

The shared library is built in-tree by ``geomjax_amd.build`` and loaded from
``geomjax_amd/_lib/libgeomjax_amd.so``.  There is no other implementation of
the hot path: if the library is missing or fails to load, every sampler call
raises ``NativeLibraryError``.
"""
import ctypes
import os
import threading

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libgeomjax_amd.so")

GJX_F64 = 0
GJX_F32 = 1
GJX_OK, GJX_EINVAL, GJX_EUNSUPPORTED, GJX_EHIP = 0, -1, -2, -3  # enum gjx_status
GJX_UTURN = {"euc": 0, "bet": 1, "riem": 2}  # stopping criteria of the NUTS entry point

_c_int32 = ctypes.c_int32
_c_double = ctypes.c_double
_c_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p


class NativeLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or returned an error."""


class FunnelParams(ctypes.Structure):
    _fields_ = [
        ("num_chains", _c_int32),
        ("dim", _c_int32),
        ("dtype", _c_int32),
        ("num_integration_steps", _c_int32),
        ("max_iters", _c_int32),
        ("num_transitions", _c_int32),
        ("lanes_per_chain", _c_int32),
        ("flags", _c_int32),
        ("step_size", _c_double),
        ("alpha2", _c_double),
        ("sigma", _c_double),
        ("convergence_tol", _c_double),
        ("divergence_tol", _c_double),
        ("divergence_threshold", _c_double),
        ("seed", _c_u64),
        ("counter", _c_u64),
    ]


class RmhmcStepIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "logdensity_grad", "z", "u",
        "new_position", "new_logdensity", "new_logdensity_grad",
        "momentum", "acceptance_rate", "is_accepted", "is_divergent", "energy",
        "proposal_position", "proposal_momentum", "proposal_velocity", "proposal_logdensity",
        "proposal_logdensity_grad", "proposal_weight", "proposal_sum_log_p_accept",
        "fp_iters", "fp_norm", "fp_success", "num_integration_steps",
    )]


class NutsIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "z", "u_direction", "u_merge", "u_step",
        "new_position", "new_logdensity", "new_logdensity_grad", "momentum", "energy", "acceptance_rate",
        "is_divergent", "is_turning", "num_trajectory_expansions", "num_integration_steps",
        "leftmost_position", "rightmost_position",
    )]


class RmhmcSampleIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "z", "u",
        "final_position", "final_logdensity", "final_logdensity_grad",
        "samples", "acceptance_rate", "num_accepted", "num_fevals", "num_integration_steps",
    )]


class LmcStepIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "inverse_mass_matrix", "position", "logdensity", "logdensity_grad", "volume_adjustment", "z", "u",
        "new_position", "new_logdensity", "new_logdensity_grad", "new_volume_adjustment",
        "velocity", "acceptance_rate", "is_accepted", "is_divergent", "energy",
        "proposal_position", "proposal_momentum", "proposal_velocity", "proposal_logdensity",
        "proposal_logdensity_grad_norm", "proposal_dl_ig", "proposal_Hdl_ig", "proposal_ig_Hdl_ig",
        "proposal_logdensity_hvp_velocity_norm", "proposal_determinant_metric", "proposal_volume_adjustment",
        "proposal_weight", "proposal_sum_log_p_accept", "num_integration_steps",
    )]


class LmcSampleIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "inverse_mass_matrix", "position", "logdensity", "volume_adjustment", "z", "u",
        "final_position", "final_logdensity", "final_logdensity_grad", "final_volume_adjustment",
        "samples", "acceptance_rate", "num_accepted", "num_integration_steps",
    )]


class LmcNutsIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "inverse_mass_matrix", "position", "logdensity", "volume_adjustment", "z", "u_direction", "u_merge", "u_step",
        "new_position", "new_logdensity", "new_logdensity_grad", "new_volume_adjustment", "velocity", "energy",
        "acceptance_rate", "is_divergent", "is_turning", "num_trajectory_expansions", "num_integration_steps",
        "leftmost_position", "rightmost_position",
    )]


class LmcAlpha2TangentIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "inverse_mass_matrix", "position", "z", "num_integration_steps", "proposal_position",
        "d_proposal_position_d_alpha2",
    )]


class LogisticParams(ctypes.Structure):
    _fields_ = [
        ("num_chains", _c_int32),
        ("dim", _c_int32),
        ("dtype", _c_int32),
        ("num_integration_steps", _c_int32),
        ("max_iters", _c_int32),
        ("num_transitions", _c_int32),
        ("flags", _c_int32),
        ("waves_per_chain", _c_int32),
        ("step_size", _c_double),
        ("prior_var", _c_double),
        ("convergence_tol", _c_double),
        ("divergence_tol", _c_double),
        ("divergence_threshold", _c_double),
        ("seed", _c_u64),
        ("counter", _c_u64),
    ]


class LogisticData(ctypes.Structure):
    _fields_ = [
        ("num_posteriors", _c_int32),
        ("max_rows", _c_int32),
        ("max_features", _c_int32),
        ("reserved", _c_int32),
    ] + [(n, _vp) for n in ("num_rows", "num_features", "x_offset", "y_offset", "chain_posterior", "x", "y",
                            "packed", "workspace")] + [("workspace_bytes", ctypes.c_uint64),
                                                       ("dispatch_order", _vp)]


class NutsSampleIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "z", "u_direction", "u_merge", "u_step", "final_position", "final_logdensity",
        "final_logdensity_grad", "samples", "acceptance_rate", "num_integration_steps", "num_fevals",
    )]


class DiagnosticsArgs(ctypes.Structure):
    _fields_ = [("num_samples", _c_int32), ("num_chains", _c_int32), ("num_coords", _c_int32), ("dtype", _c_int32),
                ("stride_sample", ctypes.c_int64), ("stride_chain", ctypes.c_int64),
                ("stride_coord", ctypes.c_int64)]


class TangentIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "z", "objective", "acceptance", "sq_jump", "d_objective", "d_acceptance",
        "d_sq_jump", "num_fevals",
    )]


class LmcmcStepIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "logdensity_grad", "volume_adjustment", "z", "u",
        "new_position", "new_logdensity", "new_logdensity_grad", "new_volume_adjustment",
        "velocity", "acceptance_rate", "is_accepted", "is_divergent", "energy",
        "proposal_position", "proposal_momentum", "proposal_velocity", "proposal_logdensity",
        "proposal_logdensity_grad", "proposal_volume_adjustment", "proposal_weight", "proposal_sum_log_p_accept",
        "num_integration_steps",
    )]


class LmcmcSampleIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "volume_adjustment", "z", "u", "final_position", "final_logdensity",
        "final_logdensity_grad", "final_volume_adjustment", "samples", "acceptance_rate", "num_accepted",
        "num_integration_steps",
    )]


class MmalaStepIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "logdensity_grad", "z", "u", "new_position", "new_logdensity",
        "new_logdensity_grad", "new_metric", "new_gamma", "acceptance_rate", "is_accepted", "is_divergent", "weight",
        "proposal_position", "proposal_logdensity", "proposal_logdensity_grad", "proposal_metric", "proposal_gamma",
    )]


class MmalaSampleIO(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "position", "logdensity", "z", "u", "final_position", "final_logdensity", "final_logdensity_grad",
        "final_metric", "final_gamma", "samples", "acceptance_rate", "num_accepted",
    )]


class DenseMetricIO(ctypes.Structure):
    _fields_ = [("num_chains", _c_int32), ("dim", _c_int32), ("step_size", _c_double)] + [(n, _vp) for n in (
        "metric", "metric_jacobian", "z", "velocity", "logdensity_grad", "grad_logdet_in", "velocity_out",
        "logdet", "grad_logdet", "vgv", "volume_adjustment", "gamma", "drift", "chol_inv_z", "momentum",
        "momentum_out", "dtdq", "pu", "step_sizes",
    )]


LMC_EXACT_SHERMAN_MORRISON = 1
LMC_EXACT_MOMENTUM = 2

# entry point name -> (argtypes) ; every entry point returns int
_SIGNATURES = {
    "gjx_rmhmc_monge_funnel_step": [ctypes.POINTER(FunnelParams), ctypes.POINTER(RmhmcStepIO), _vp],
    "gjx_rmhmc_monge_funnel_sample": [ctypes.POINTER(FunnelParams), ctypes.POINTER(RmhmcSampleIO), _vp],
    "gjx_lmcmonge_funnel_step": [ctypes.POINTER(FunnelParams), ctypes.POINTER(LmcStepIO), _vp],
    "gjx_lmcmonge_funnel_sample": [ctypes.POINTER(FunnelParams), ctypes.POINTER(LmcSampleIO), _vp],
    "gjx_lmcmonge_funnel_alpha2_tangent": [ctypes.POINTER(FunnelParams), ctypes.POINTER(LmcAlpha2TangentIO), _vp],
    "gjx_lmcmonge_funnel_nuts": [ctypes.POINTER(FunnelParams), _c_int32, _c_int32, ctypes.POINTER(LmcNutsIO), _vp],
    "gjx_lmc_fisher_logistic_nuts": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _c_int32, _c_int32,
                                     ctypes.POINTER(LmcNutsIO), _vp],
    "gjx_rmhmc_monge_funnel_tangent": [ctypes.POINTER(FunnelParams), ctypes.POINTER(TangentIO), _vp],
    "gjx_rmhmc_monge_funnel_nuts": [ctypes.POINTER(FunnelParams), _c_int32, _c_int32, ctypes.POINTER(NutsIO), _vp],
    "gjx_rmhmc_fisher_logistic_nuts": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _c_int32,
                                       _c_int32, ctypes.POINTER(NutsIO), _vp],
    "gjx_funnel_value_and_grad": [_c_int32, _c_int32, _c_double, _c_int32, _vp, _vp, _vp, _vp],
    "gjx_rmhmc_fisher_logistic_step": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                       ctypes.POINTER(RmhmcStepIO), _vp],
    "gjx_rmhmc_fisher_logistic_step_noalloc": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                               ctypes.POINTER(RmhmcStepIO), _vp],
    "gjx_rmhmc_fisher_logistic_sample": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                         ctypes.POINTER(RmhmcSampleIO), _vp],
    "gjx_lmc_fisher_logistic_step": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                     ctypes.POINTER(LmcmcStepIO), _vp],
    "gjx_lmc_fisher_logistic_sample": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                       ctypes.POINTER(LmcmcSampleIO), _vp],
    "gjx_logistic_value_and_grad": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _vp, _vp, _vp,
                                    _vp],
    "gjx_dense_metric_position": [ctypes.POINTER(DenseMetricIO), _vp],
    "gjx_dense_lmc_half_step": [ctypes.POINTER(DenseMetricIO), _vp],
    "gjx_dense_mmala_drift": [ctypes.POINTER(DenseMetricIO), _vp],
    "gjx_dense_rmhmc_field": [ctypes.POINTER(DenseMetricIO), _vp],
    "gjx_mmala_fisher_logistic_init": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _vp, _vp, _vp,
                                       _vp, _vp, _vp],
    "gjx_mmala_fisher_logistic_step": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _c_int32,
                                       ctypes.POINTER(MmalaStepIO), _vp],
    "gjx_mmala_fisher_logistic_sample": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData), _c_int32,
                                         ctypes.POINTER(MmalaSampleIO), _vp],
    "gjx_rmhmc_fisher_logistic_tangent": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                          ctypes.POINTER(TangentIO), _vp],
    "gjx_logistic_pack_size": [_c_int32, _vp, _vp, ctypes.POINTER(ctypes.c_uint64)],
    "gjx_logistic_pack": [ctypes.POINTER(LogisticData), _vp, _vp, _vp, ctypes.c_uint64, _vp],
    "gjx_logistic_workspace_size": [ctypes.POINTER(LogisticParams), ctypes.POINTER(LogisticData),
                                    ctypes.POINTER(ctypes.c_uint64)],
}
_SIGNATURES["gjx_rmhmc_monge_funnel_nuts_sample"] = [ctypes.POINTER(FunnelParams), _c_int32, _c_int32,
                                                     ctypes.POINTER(NutsSampleIO), _vp]
_SIGNATURES["gjx_diagnostics_workspace_size"] = [ctypes.POINTER(DiagnosticsArgs), ctypes.POINTER(ctypes.c_uint64)]
_SIGNATURES["gjx_effective_sample_size"] = [ctypes.POINTER(DiagnosticsArgs), _vp, _vp, ctypes.c_uint64, _vp, _vp, _vp]

_STRING_FUNCS = ("gjx_version", "gjx_last_error")

_lib = None
_lock = threading.Lock()


def load(path=None):
    """Load (once) and return the ctypes handle of the native library."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise NativeLibraryError(
                f"{p} does not exist: build it with `python -m geomjax_amd.build` "
                "(there is no non-HIP implementation of the sampler kernels)"
            )
        try:
            lib = ctypes.CDLL(p)
        except OSError as e:
            raise NativeLibraryError(f"cannot load {p}: {e}") from e
        for name, argtypes in _SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = argtypes
            f.restype = ctypes.c_int
        for name in _STRING_FUNCS:
            f = getattr(lib, name)
            f.argtypes = []
            f.restype = ctypes.c_char_p
        _lib = lib
        return lib


def exported_symbols():
    return list(_SIGNATURES) + list(_STRING_FUNCS)


def last_error():
    return load().gjx_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise NativeLibraryError(f"{what} failed with status {rc}: {last_error()}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL).  A host tensor raises: its
    address is not a device address, and a kernel reading it would fault."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"expected a tensor on a HIP device, got a {t.dtype} tensor on {t.device}")
    return t.data_ptr()


def stream_handle(device):
    """hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream
