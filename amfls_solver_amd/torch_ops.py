"""torch.ops.amfls registration on the HIP engine.

Registers the reference's schemas for run_amfls, run_fixed_rbgk and run_lsqr
(reference bindings/torch/torch_binding.cpp:510-525) with implementations
for dense and sparse-CSR tensors on the CPU and on the GPU
(torch_binding.cpp:581-611 registers CPU and SparseCsrCPU only).  Each call
runs the HIP engine (host tensors are copied to the device first) and
returns the reference's 8-tuple (solution, metrics[21], counters[36],
timings[9], trace_float[c, 29], trace_int[c, 4], status, stop_reason) as CPU
tensors, bindings/torch/result_tensors.hpp:12-156.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .binding import _UINT64_MODULUS, _operator, _result_tuple, _rhs, run_amfls_raw

SCHEMAS = {
    "run_amfls": (
        "run_amfls(Tensor matrix, Tensor b, float regularization, float tolerance, "
        "float failure_probability, int maximum_epochs=0, int maximum_depth=0, "
        "int maximum_auxiliary_width=0, int maximum_basis_size=0, "
        "int seed=0, int stream=0) -> "
        "(Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)"),
    "run_fixed_rbgk": (
        "run_fixed_rbgk(Tensor matrix, Tensor b, float regularization, float tolerance, "
        "float failure_probability, int auxiliary_width, int maximum_depth, "
        "int maximum_basis_size, int seed, int stream) -> "
        "(Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)"),
    "run_lsqr": (
        "run_lsqr(Tensor matrix, Tensor b, float regularization, float tolerance, "
        "int maximum_iterations=0) -> "
        "(Tensor, Tensor, Tensor, Tensor, Tensor, Tensor, int, int)"),
}

DISPATCH_KEYS = ("CPU", "SparseCsrCPU", "CUDA", "SparseCsrCUDA")


def _run_amfls(matrix, b, regularization, tolerance, failure_probability, maximum_epochs=0, maximum_depth=0,
               maximum_auxiliary_width=0, maximum_basis_size=0, seed=0, stream=0):
    raw, _, _ = run_amfls_raw(matrix, b, regularization, tolerance, failure_probability, maximum_epochs,
                              maximum_depth, maximum_auxiliary_width, maximum_basis_size, seed, stream)
    return raw


def _run_fixed_rbgk(matrix, b, regularization, tolerance, failure_probability, auxiliary_width, maximum_depth,
                    maximum_basis_size, seed, stream):
    op = _operator(matrix, None, None)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_fixed_rbgk(ctx.handle, op.handle, bd.data_ptr(), 1, float(regularization),
                                               float(tolerance), float(failure_probability), int(auxiliary_width),
                                               int(maximum_depth), int(maximum_basis_size),
                                               int(seed) % _UINT64_MODULUS, int(stream) % _UINT64_MODULUS,
                                               _lib.CLOCK["wall"], ctypes.byref(res)))
    return _result_tuple(ctx, res)[0]


def _run_lsqr(matrix, b, regularization, tolerance, maximum_iterations=0):
    op = _operator(matrix, None, None)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_lsqr(ctx.handle, op.handle, bd.data_ptr(), 1, float(regularization),
                                         float(tolerance), int(maximum_iterations), _lib.CLOCK["wall"],
                                         ctypes.byref(res)))
    return _result_tuple(ctx, res)[0]


IMPLS = {"run_amfls": _run_amfls, "run_fixed_rbgk": _run_fixed_rbgk, "run_lsqr": _run_lsqr}

_library = None


def register() -> None:
    """Define the amfls ops (once per process) and bind them to the engine."""
    global _library
    if _library is not None:
        return
    try:
        torch.ops.amfls.run_amfls  # noqa: B018  (another provider already defined the namespace)
        return
    except (AttributeError, RuntimeError):
        pass
    lib = torch.library.Library("amfls", "DEF")
    for name, schema in SCHEMAS.items():
        lib.define(schema)
        for key in DISPATCH_KEYS:
            lib.impl(name, IMPLS[name], key)
    _library = lib
