"""CPU checks of the C ABI boundary: the library loads and exports every
function include/amfls_hip.h declares (no GPU calls)."""
import ctypes

from amfls_solver_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 25
    missing = [name for name in declared if not hasattr(lib, name)]
    assert missing == []
    # every declared entry point has a ctypes signature in the loader
    assert sorted(_lib.SIGNATURES) == declared


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.amfls_version() >= 10000
    o = _lib.options()
    assert o.tolerance == 1e-8 and o.failure_probability == 1e-6 and o.precision == 0
    n = lib.amfls_spmm_variant_count()
    assert n >= 8
    names = {lib.amfls_spmm_variant_name(i).decode() for i in range(n)}
    assert len(names) == n
    assert lib.amfls_spmm_variant_supports(0, 1) == 1
    assert lib.amfls_spmm_variant_supports(0, 2) == 0


def test_errors_cross_the_boundary_as_codes():
    lib = _lib.load()
    assert lib.amfls_context_create(None) == 1  # invalid argument, no exception
    assert "null" in lib.amfls_last_error().decode()
    assert lib.amfls_operator_shape(None, None, None) == 1
    res = ctypes.c_void_p()
    assert lib.amfls_solve(None, None, None, 0, None, ctypes.byref(res)) == 1
