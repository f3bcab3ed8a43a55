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
