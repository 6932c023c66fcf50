"""ctypes loader for libamfls_hip.so (the C ABI declared in include/amfls_hip.h).

torch is imported before the library so that the process uses a single HIP
runtime (the one torch loaded); the library then resolves its HIP symbols
against it.  A missing library is an error: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libamfls_hip.so"

F64, F32, BF16 = 0, 1, 2
PRECISION = {"fp64": 0, "mixed": 1, "fp32": 2}
WIDTH_POLICY = {"heuristic": 0, "learned": 1, "record": 2, "widen": 3, "hold": 4, "record_widen": 5}
WIDTH_INPUTS, WIDTH_PARAMS = 12, 5057
CLOCK = {"wall": 0, "synthetic": 1}

ERRORS = {
    1: ValueError,      # invalid_argument
    2: RuntimeError,    # logic_error
    3: OverflowError,   # length_error
    4: RuntimeError,    # runtime_error
    5: RuntimeError,    # HIP failure
    6: RuntimeError,
}


class Options(ctypes.Structure):
    _fields_ = [
        ("regularization", c_double),
        ("tolerance", c_double),
        ("failure_probability", c_double),
        ("maximum_epochs", c_int32),
        ("maximum_basis_size", c_int32),
        ("seed", c_uint64),
        ("stream", c_uint64),
        ("maximum_depth", c_int32),
        ("maximum_auxiliary_width", c_int32),
        ("precision", c_int32),
        ("clock_mode", c_int32),
        ("synthetic_operator_unit", c_double),
        ("synthetic_local_unit", c_double),
        ("maximum_refinements", c_int32),
        ("inner_tolerance", c_double),
        ("width_policy", c_int32),
        ("forward_tolerance", c_double),
        ("maximum_refinement_iterations", c_int32),
        ("width_script", c_void_p),
        ("width_script_length", c_int32),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "amfls_last_error": (c_char_p, []),
    "amfls_version": (c_int, []),
    "amfls_options_default": (None, [POINTER(Options)]),
    "amfls_context_create": (c_int, [POINTER(c_void_p)]),
    "amfls_context_destroy": (c_int, [c_void_p]),
    "amfls_context_synchronize": (c_int, [c_void_p]),
    "amfls_csr_create": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int,
                                 POINTER(c_void_p)]),
    "amfls_dense_create": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_int, c_int, c_int64,
                                   POINTER(c_void_p)]),
    "amfls_host_operator_create": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_double,
                                           c_double, c_int64, c_int64, POINTER(c_void_p)]),
    "amfls_host_operator_create_ex": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_double, c_double, c_int64,
                                              c_int64, POINTER(c_void_p)]),
    "amfls_operator_destroy": (c_int, [c_void_p]),
    "amfls_operator_shape": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "amfls_operator_validation_model": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double),
                                                POINTER(c_int64), POINTER(c_int64)]),
    "amfls_operator_relative_block_product_cost": (c_double, [c_void_p, c_int]),
    "amfls_operator_apply": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int]),
    "amfls_operator_apply_rows": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int64, c_void_p,
                                          c_int64, c_int]),
    "amfls_operator_apply_rows_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int64, c_void_p,
                                                  c_int64, c_int]),
    "amfls_operator_product_bytes": (c_double, [c_void_p, c_int, c_int, c_int]),
    "amfls_operator_product_flops": (c_double, [c_void_p, c_int]),
    "amfls_operator_set_variant": (c_int, [c_void_p, c_int, c_int, c_int]),
    "amfls_operator_profile": (c_int, [c_void_p, c_int]),
    "amfls_operator_kernel_stats": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_double), POINTER(c_double)]),
    "amfls_operator_kernel_stats_class": (c_int, [c_void_p, c_int, c_int32, c_int32, c_int, POINTER(c_int64),
                                                  POINTER(c_double), POINTER(c_double)]),
    "amfls_spmm_variant_count": (c_int, []),
    "amfls_spmm_variant_name": (c_char_p, [c_int]),
    "amfls_spmm_variant_supports": (c_int, [c_int, c_int]),
    "amfls_spmm_default_variant": (c_int, [c_int, c_double]),
    "amfls_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(Options), POINTER(c_void_p)]),
    "amfls_solve_fixed_rbgk": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_double,
                                       c_int32, c_int32, c_int32, c_uint64, c_uint64, c_int32,
                                       POINTER(c_void_p)]),
    "amfls_solve_lsqr": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_int32, c_int32,
                                 POINTER(c_void_p)]),
    "amfls_validate_candidate": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_double, c_double,
                                         POINTER(c_void_p)]),
    "amfls_block_rank_reveal": (c_int, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_int, c_int32, c_double,
                                        POINTER(c_int32), POINTER(c_double), POINTER(c_double), c_void_p]),
    "amfls_result_sizes": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "amfls_result_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  POINTER(c_int64), POINTER(c_int64)]),
    "amfls_result_trace_extension": (c_int, [c_void_p, c_void_p]),
    "amfls_result_refinement_steps": (c_int, [c_void_p, POINTER(c_int32)]),
    "amfls_result_forward_error_estimate": (c_int, [c_void_p, POINTER(c_double)]),
    "amfls_result_fp64_fallback": (c_int, [c_void_p, POINTER(c_int32)]),
    "amfls_result_destroy": (c_int, [c_void_p]),
    "amfls_gaussian_fill": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int64, c_int, c_uint64, c_uint64,
                                    c_uint64]),
    "amfls_result_width_records": (c_int, [c_void_p, c_void_p, POINTER(c_int64)]),
    "amfls_block_gemm_nn": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_int64,
                                    c_void_p, c_int64, c_double, c_double]),
    "amfls_block_gemm_tn": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_int64,
                                    c_void_p, c_int64, c_double]),
    "amfls_block_rowmul": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_void_p,
                                   c_int64]),
    "amfls_block_gemm_nn_f32": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_int64,
                                        c_void_p, c_int64, c_double, c_double]),
    "amfls_block_gemm_tn_f32": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_int64,
                                        c_void_p, c_int64, c_double]),
    "amfls_block_rowmul_f32": (c_int, [c_void_p, c_int64, c_int32, c_int32, c_void_p, c_int64, c_void_p, c_void_p,
                                       c_int64]),
    "amfls_block_stats": (c_int, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p]),
    "amfls_width_model_create": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "amfls_width_model_destroy": (c_int, [c_void_p]),
    "amfls_width_model_get_params": (c_int, [c_void_p, c_void_p]),
    "amfls_width_model_set_params": (c_int, [c_void_p, c_void_p]),
    "amfls_width_model_logits": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "amfls_width_model_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, POINTER(c_float)]),
    "amfls_width_model_train_step": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_float, POINTER(c_float)]),
    "amfls_csr_generate_rows": (c_int, [c_void_p, c_int64, c_int64, c_int32, c_int32, c_uint64, c_void_p,
                                        POINTER(c_int64)]),
    "amfls_csr_generate_entries": (c_int, [c_void_p, c_int64, c_int64, c_uint64, c_double, c_double, c_void_p,
                                           c_void_p, c_void_p, c_int]),
    "amfls_operator_autotune": (c_int, [c_void_p, c_void_p, POINTER(c_int32), c_int32, c_int, c_int32,
                                        POINTER(c_int32)]),
    "amfls_solve_sweep": (c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(c_double), c_int32,
                                  POINTER(Options), POINTER(c_void_p)]),
    "amfls_solve_batch": (c_int, [c_void_p, c_int32, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_double),
                                  c_int32, POINTER(Options), c_int32, POINTER(c_void_p)]),
    "amfls_width_model_set_domain": (c_int, [c_void_p, c_void_p, c_void_p]),
    "amfls_width_model_decisions": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    "amfls_context_set_width_model": (c_int, [c_void_p, c_void_p]),
    "amfls_context_set_level_graphs": (c_int, [c_void_p, c_int]),
}

_lib = None


def declared_symbols() -> list[str]:
    """Function names declared in include/amfls_hip.h."""
    import re
    header = Path(__file__).resolve().parent.parent / "include" / "amfls_hip.h"
    text = header.read_text()
    return sorted(set(re.findall(r"\b(amfls_[a-z0-9_]+)\s*\(", text)))


def load(path: Path | None = None):
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is part of this image
        pass
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m amfls_solver_amd.build`")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().amfls_last_error().decode(errors="replace")
        raise ERRORS.get(status, RuntimeError)(msg)


def options(**kw) -> Options:
    o = Options()
    load().amfls_options_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


__all__ = ["load", "check", "options", "Options", "F64", "F32", "BF16", "PRECISION", "CLOCK",
           "declared_symbols", "c_float"]
