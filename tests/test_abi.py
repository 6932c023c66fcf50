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


def test_torch_ops_schemas_match_reference():
    """torch.ops.amfls carries the reference's schemas
    (bindings/torch/torch_binding.cpp:510-525)."""
    import torch
    import amfls_solver_amd  # noqa: F401  (registers the ops)
    want = {
        "run_amfls": "amfls::run_amfls(Tensor matrix, Tensor b, float regularization, float tolerance, "
                     "float failure_probability, int maximum_epochs=0, int maximum_depth=0, "
                     "int maximum_auxiliary_width=0, int maximum_basis_size=0, int seed=0, int stream=0) -> "
                     "(Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)",
        "run_fixed_rbgk": "amfls::run_fixed_rbgk(Tensor matrix, Tensor b, float regularization, float tolerance, "
                          "float failure_probability, int auxiliary_width, int maximum_depth, "
                          "int maximum_basis_size, int seed, int stream) -> "
                          "(Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)",
        "run_lsqr": "amfls::run_lsqr(Tensor matrix, Tensor b, float regularization, float tolerance, "
                    "int maximum_iterations=0) -> (Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)",
    }
    for name, schema in want.items():
        assert str(getattr(torch.ops.amfls, name).default._schema) == schema
