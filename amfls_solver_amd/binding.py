"""Python mirror of the reference's AMFLS binding on the MI355X engine.

Same entry points, argument meaning and result type as the reference's
python/amfls/binding.py (run_amfls :259-290, run_fixed_rbgk, run_lsqr,
SolveResult :9-90, _decode :159-255) and the same 8-tuple layout as
bindings/torch/result_tensors.hpp:12-156.  Matrices may be CPU or CUDA
(HIP) tensors, dense strided or sparse CSR; they are moved to the GPU, the
solve runs in libamfls_hip.so, and results come back as CPU tensors like the
reference's.

Engine extensions (keyword-only): ``precision`` ("fp64" or "mixed"),
``clock`` ("wall" or "synthetic"), ``value_dtype`` (storage of A on the
device: torch.float64, torch.float32 or torch.bfloat16 for dense A).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields

import numpy as np
import torch

from . import _lib

_UINT64_MODULUS = 1 << 64
_INT64_SIGN_BIT = 1 << 63


@dataclass(frozen=True)
class SolveResult:
    solution: torch.Tensor
    objective: float
    residual_norm: float
    augmented_residual_norm: float
    normal_residual_norm: float
    solution_norm: float
    solution_energy_norm: float
    operator_norm_lower_bound: float
    augmented_operator_norm_lower_bound: float
    compatible_backward_error_upper_bound: float
    least_squares_backward_error_upper_bound: float
    backward_error_upper_bound: float
    relative_normal_residual_upper_bound: float
    energy_error_upper_bound: float
    relative_energy_error_upper_bound: float
    objective_gap_upper_bound: float
    ridge_base_energy_error_upper_bound: float | None
    ridge_corrected_energy_error_upper_bound: float | None
    ridge_correction_gamma: float | None
    ridge_correction_z_h_z: float | None
    ridge_correction_two_abs_z_t_q: float | None
    ridge_correction_q_norm_squared: float | None
    a_columns: int
    at_columns: int
    a_block_calls: int
    at_block_calls: int
    search_a_columns: int
    search_at_columns: int
    validation_a_columns: int
    validation_at_columns: int
    iterations: int
    depth: int
    auxiliary_width: int
    basis_rank: int
    sketch_a_columns: int
    sketch_at_columns: int
    sketch_a_block_calls: int
    sketch_at_block_calls: int
    iterative_a_columns: int
    iterative_at_columns: int
    iterative_a_block_calls: int
    iterative_at_block_calls: int
    gaussian_random_columns: int
    gaussian_random_values: int
    search_a_block_calls: int
    search_at_block_calls: int
    validation_a_block_calls: int
    validation_at_block_calls: int
    base_validation_a_columns: int
    base_validation_at_columns: int
    base_validation_a_block_calls: int
    base_validation_at_block_calls: int
    ridge_correction_a_columns: int
    ridge_correction_at_columns: int
    ridge_correction_a_block_calls: int
    ridge_correction_at_block_calls: int
    gaussian_random_block_requests: int
    ridge_correction_disposition: int
    total_seconds: float
    a_seconds: float
    at_seconds: float
    orthogonalization_seconds: float
    projected_solve_seconds: float
    validation_seconds: float
    other_seconds: float
    base_validation_seconds: float
    ridge_correction_validation_seconds: float
    trace: torch.Tensor
    trace_float: torch.Tensor
    trace_int: torch.Tensor
    status: int
    stop_reason: int
    # engine extensions
    trace_extension: torch.Tensor | None = None
    refinement_steps: int = 0
    # mixed precision: forward-error estimate ||P P^T g|| / ||x|| (NaN if none)
    forward_error_estimate: float = float("nan")
    fp64_fallback: bool = False

    @property
    def success(self) -> bool:
        return self.status == 0

    @property
    def search_vector_work(self) -> int:
        return self.search_a_columns + self.search_at_columns


_METRIC_NAMES = [f.name for f in fields(SolveResult)][1:22]
_COUNTER_NAMES = [
    "a_columns", "at_columns", "a_block_calls", "at_block_calls", "search_a_columns", "search_at_columns",
    "validation_a_columns", "validation_at_columns", "iterations", "depth", "auxiliary_width", "basis_rank",
    "sketch_a_columns", "sketch_at_columns", "sketch_a_block_calls", "sketch_at_block_calls",
    "iterative_a_columns", "iterative_at_columns", "iterative_a_block_calls", "iterative_at_block_calls",
    "gaussian_random_columns", "gaussian_random_values", "search_a_block_calls", "search_at_block_calls",
    "validation_a_block_calls", "validation_at_block_calls", "base_validation_a_columns",
    "base_validation_at_columns", "base_validation_a_block_calls", "base_validation_at_block_calls",
    "ridge_correction_a_columns", "ridge_correction_at_columns", "ridge_correction_a_block_calls",
    "ridge_correction_at_block_calls", "gaussian_random_block_requests", "ridge_correction_disposition",
]
_TIMING_NAMES = ["total_seconds", "a_seconds", "at_seconds", "orthogonalization_seconds",
                 "projected_solve_seconds", "validation_seconds", "other_seconds", "base_validation_seconds",
                 "ridge_correction_validation_seconds"]


class Context:
    """One engine context (HIP stream + scratch) on the current device."""

    def __init__(self):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.amfls_context_create(ctypes.byref(h)))
        self.handle = h

    def synchronize(self) -> None:
        _lib.check(self._lib.amfls_context_synchronize(self.handle))

    def set_level_graphs(self, enabled: bool) -> None:
        """Capture width-one level batches into HIP graphs (True) or enqueue
        them kernel by kernel (False, the default); results are identical."""
        _lib.check(self._lib.amfls_context_set_level_graphs(self.handle, int(bool(enabled))))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._lib.amfls_context_destroy(h)
            self.handle = None


_default_context: Context | None = None


def default_context() -> Context:
    global _default_context
    if _default_context is None:
        _default_context = Context()
    return _default_context


def _require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the AMFLS HIP engine needs a ROCm GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


_DTYPE_CODE = {torch.float64: _lib.F64, torch.float32: _lib.F32, torch.bfloat16: _lib.BF16}


class DeviceMatrix:
    """A device operator built from a torch tensor (dense strided or sparse CSR).

    The reference builds a RowMajorDenseOperator from contiguous dense tensors
    and a SparseCsrOperator from CSR tensors (torch_binding.cpp:69-114); this
    keeps the device copies alive for the operator's lifetime.
    """

    def __init__(self, matrix: torch.Tensor, *, value_dtype: torch.dtype | None = None,
                 context: Context | None = None):
        dev = _require_gpu()
        self.context = context or default_context()
        lib = self.context._lib
        if matrix.ndim != 2 or matrix.shape[0] <= 0 or matrix.shape[1] <= 0:
            raise ValueError("matrix must have shape (m,n) with m, n > 0")
        if matrix.shape[0] > 2**31 - 1 or matrix.shape[1] > 2**31 - 1:
            raise ValueError("matrix exceeds the LP64 dimension limit")
        self.rows, self.cols = int(matrix.shape[0]), int(matrix.shape[1])
        h = ctypes.c_void_p()
        if matrix.layout == torch.sparse_csr:
            vd = value_dtype or (matrix.dtype if matrix.dtype in (torch.float64, torch.float32) else torch.float64)
            if vd not in (torch.float64, torch.float32):
                raise ValueError("CSR values are stored as float64 or float32")
            self._rp = matrix.crow_indices().to(device=dev, dtype=torch.int64).contiguous()
            self._ci = matrix.col_indices().to(device=dev, dtype=torch.int32).contiguous()
            self._val = matrix.values().to(device=dev, dtype=vd).contiguous()
            if self._rp.numel() != self.rows + 1 or self._ci.numel() != self._val.numel():
                raise ValueError("CSR row offsets must have length rows + 1")
            torch.cuda.synchronize()
            _lib.check(lib.amfls_csr_create(self.context.handle, self.rows, self.cols, self._val.numel(),
                                            self._rp.data_ptr(), self._ci.data_ptr(), self._val.data_ptr(),
                                            _DTYPE_CODE[vd], ctypes.byref(h)))
            self.kind = "csr"
            self.nnz = int(self._val.numel())
        elif matrix.layout == torch.strided:
            vd = value_dtype or (matrix.dtype if matrix.dtype in _DTYPE_CODE else torch.float64)
            if vd not in _DTYPE_CODE:
                raise ValueError("dense values are stored as float64, float32 or bfloat16")
            t = matrix.to(device=dev, dtype=vd)
            if t.stride(1) == 1 and t.stride(0) >= self.cols:
                self._val, row_major, ld = t, 1, t.stride(0)
            elif t.stride(0) == 1 and t.stride(1) >= self.rows:
                self._val, row_major, ld = t, 0, t.stride(1)
            else:
                self._val, row_major, ld = t.contiguous(), 1, self.cols
            torch.cuda.synchronize()
            _lib.check(lib.amfls_dense_create(self.context.handle, self.rows, self.cols, self._val.data_ptr(),
                                              _DTYPE_CODE[vd], row_major, ld, ctypes.byref(h)))
            self.kind = "dense"
            self.nnz = self.rows * self.cols
        else:
            raise ValueError("matrix must use dense strided or sparse CSR storage")
        self.handle = h
        self.value_dtype = vd

    @classmethod
    def from_csr_arrays(cls, rows: int, cols: int, row_ptr: torch.Tensor, col_idx: torch.Tensor,
                        values: torch.Tensor, *, context: Context | None = None) -> "DeviceMatrix":
        """CSR operator over device arrays (int64 row pointers, int32 column
        indices, float32/float64 values) that the caller keeps unchanged."""
        dev = _require_gpu()
        self = cls.__new__(cls)
        self.context = context or default_context()
        if rows <= 0 or cols <= 0 or rows > 2**31 - 1 or cols > 2**31 - 1:
            raise ValueError("matrix must have shape (m,n) with 0 < m, n < 2^31")
        if row_ptr.dtype != torch.int64 or col_idx.dtype != torch.int32 or values.dtype not in (torch.float64,
                                                                                                torch.float32):
            raise ValueError("CSR arrays must be int64 row pointers, int32 columns and float32/float64 values")
        if row_ptr.numel() != rows + 1 or col_idx.numel() != values.numel():
            raise ValueError("CSR row offsets must have length rows + 1")
        self.rows, self.cols = int(rows), int(cols)
        self._rp = row_ptr.to(device=dev).contiguous()
        self._ci = col_idx.to(device=dev).contiguous()
        self._val = values.to(device=dev).contiguous()
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        _lib.check(self.context._lib.amfls_csr_create(self.context.handle, self.rows, self.cols, self._val.numel(),
                                                      self._rp.data_ptr(), self._ci.data_ptr(),
                                                      self._val.data_ptr(), _DTYPE_CODE[values.dtype],
                                                      ctypes.byref(h)))
        self.kind, self.nnz, self.handle, self.value_dtype = "csr", int(self._val.numel()), h, values.dtype
        return self

    def validation_model(self):
        lib = self.context._lib
        up, lo = ctypes.c_double(), ctypes.c_double()
        sa, st = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.amfls_operator_validation_model(self.handle, ctypes.byref(up), ctypes.byref(lo),
                                                       ctypes.byref(sa), ctypes.byref(st)))
        return up.value, lo.value, sa.value, st.value

    def relative_block_product_cost(self, s: int) -> float:
        return self.context._lib.amfls_operator_relative_block_product_cost(self.handle, s)

    def apply(self, x: torch.Tensor, transpose: bool = False) -> torch.Tensor:
        """Column-major block product (MatrixOperator::apply contract)."""
        if x.ndim == 1:
            x = x[:, None]
        if x.dtype not in (torch.float64, torch.float32):
            raise ValueError("blocks must be float64 or float32")
        rows_in = self.rows if transpose else self.cols
        rows_out = self.cols if transpose else self.rows
        if x.shape[0] != rows_in:
            raise ValueError("block has the wrong number of rows")
        s = int(x.shape[1])
        xin = x.to(device=self._val.device).t().contiguous().t()  # column-major storage
        out = torch.empty((s, rows_out), dtype=x.dtype, device=xin.device).t()
        torch.cuda.synchronize()
        _lib.check(self.context._lib.amfls_operator_apply(self.context.handle, self.handle, int(transpose),
                                                          xin.data_ptr(), s, out.data_ptr(),
                                                          _DTYPE_CODE[x.dtype]))
        return out

    def set_variant(self, s: int, transpose: bool, variant: int) -> None:
        _lib.check(self.context._lib.amfls_operator_set_variant(self.handle, s, int(transpose), variant))

    def autotune(self, widths=(1, 2, 4, 8, 16, 32, 64), dtype: torch.dtype = torch.float64, reps: int = 5):
        """Time the SpMM kernel variants for each block width on this CSR
        matrix and keep the fastest for A and A^T; returns
        {width: (A variant name, A^T variant name)}."""
        w = (ctypes.c_int32 * len(widths))(*[int(x) for x in widths])
        chosen = (ctypes.c_int32 * (2 * len(widths)))()
        torch.cuda.synchronize()
        _lib.check(self.context._lib.amfls_operator_autotune(self.context.handle, self.handle, w, len(widths),
                                                             _DTYPE_CODE[dtype], int(reps), chosen))
        name = self.context._lib.amfls_spmm_variant_name
        return {int(s): (name(chosen[2 * i]).decode(), name(chosen[2 * i + 1]).decode())
                for i, s in enumerate(widths)}

    def profile(self, enable: bool = True) -> None:
        _lib.check(self.context._lib.amfls_operator_profile(self.handle, int(enable)))

    def kernel_stats(self):
        """(launches, seconds, algorithmic bytes) of timed products; resets."""
        l, sec, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self.context._lib.amfls_operator_kernel_stats(self.handle, ctypes.byref(l), ctypes.byref(sec),
                                                                ctypes.byref(by)))
        return l.value, sec.value, by.value

    def kernel_stats_class(self, transpose: bool, min_cols: int, max_cols: int,
                           dtype: torch.dtype = torch.float32):
        """(launches, seconds, algorithmic bytes) of the timed products of one
        class (direction, widths in [min_cols, max_cols], block dtype) since
        the last kernel_stats; does not reset."""
        l, sec, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self.context._lib.amfls_operator_kernel_stats_class(
            self.handle, int(transpose), min_cols, max_cols, _DTYPE_CODE[dtype], ctypes.byref(l),
            ctypes.byref(sec), ctypes.byref(by)))
        return l.value, sec.value, by.value

    def product_bytes(self, s: int, transpose: bool = False, dtype: torch.dtype = torch.float32) -> float:
        return self.context._lib.amfls_operator_product_bytes(self.handle, s, int(transpose), _DTYPE_CODE[dtype])

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.context._lib.amfls_operator_destroy(h)
            self.handle = None


_HOST_APPLY = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double))


class HostOperator(DeviceMatrix):
    """A matrix-free operator on the host: the counterpart of a custom
    amfls::MatrixOperator (reference matrix_operator.hpp:29-89).

    apply(x) gets a (cols, s) float64 array (column-major block) and returns
    A x as a (rows, s) array; apply_transpose(y) returns A^T y.  Each product
    goes through page-locked host memory.  The validation error model
    (norm_upper, norm_lower, steps_apply, steps_transpose) is the caller's;
    the default norm_upper = inf disables certification, as in the reference
    (matrix_operator.hpp:57-60)."""

    def __init__(self, rows: int, cols: int, apply, apply_transpose, *, norm_upper: float = float("inf"),
                 norm_lower: float = 0.0, steps_apply: int = 0, steps_transpose: int = 0,
                 context: Context | None = None):
        dev = _require_gpu()
        self.context = context or default_context()
        self.rows, self.cols = int(rows), int(cols)
        self._errors: list[BaseException] = []

        def wrap(fn, rin, rout):
            def cb(_user, inp, s, out):
                try:
                    x = np.ctypeslib.as_array(inp, shape=(s * rin,)).reshape((rin, s), order="F")
                    y = np.asarray(fn(x), dtype=np.float64)
                    if y.shape != (rout, s):
                        raise ValueError(f"operator callback returned shape {y.shape}, expected {(rout, s)}")
                    np.ctypeslib.as_array(out, shape=(s * rout,))[:] = y.ravel(order="F")
                    return 0
                except BaseException as e:  # reported by the engine as a failed callback
                    self._errors.append(e)
                    return 1
            return _HOST_APPLY(cb)

        self._cb_apply = wrap(apply, self.cols, self.rows)
        self._cb_apply_t = wrap(apply_transpose, self.rows, self.cols)
        self._val = torch.empty(0, device=dev)
        h = ctypes.c_void_p()
        _lib.check(self.context._lib.amfls_host_operator_create(
            self.context.handle, self.rows, self.cols, ctypes.cast(self._cb_apply, ctypes.c_void_p),
            ctypes.cast(self._cb_apply_t, ctypes.c_void_p), None, float(norm_upper), float(norm_lower),
            int(steps_apply), int(steps_transpose), ctypes.byref(h)))
        self.handle, self.kind, self.nnz, self.value_dtype = h, "host", 0, torch.float64


def generate_csr(rows: int, cols: int, nnz_range=(4, 32), *, seed: int = 0, scale_range=(1.0, 1.0),
                 value_dtype: torch.dtype = torch.float32, context: "Context | None" = None):
    """Counter-addressed synthetic CSR built on the device: per-row lengths
    uniform in nnz_range (at most 64), ascending distinct uniform columns,
    N(0,1) values times log-uniform row scales in scale_range.  Returns
    (row_ptr int64, col_idx int32, values) device tensors."""
    dev = _require_gpu()
    ctx = context or default_context()
    lo, hi = int(nnz_range[0]), int(nnz_range[1])
    rp = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    nnz = ctypes.c_int64()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_csr_generate_rows(ctx.handle, int(rows), int(cols), lo, hi, int(seed) % _UINT64_MODULUS,
                                                rp.data_ptr(), ctypes.byref(nnz)))
    ci = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)[:nnz.value]
    val = torch.empty(max(nnz.value, 1), dtype=value_dtype, device=dev)[:nnz.value]
    _lib.check(ctx._lib.amfls_csr_generate_entries(ctx.handle, int(rows), int(cols), int(seed) % _UINT64_MODULUS,
                                                   float(scale_range[0]), float(scale_range[1]), rp.data_ptr(),
                                                   ci.data_ptr(), val.data_ptr(), _DTYPE_CODE[value_dtype]))
    return rp, ci, val


def gaussian_block(rows: int, cols: int, seed: int, stream: int, *, dtype: torch.dtype = torch.float64,
                   first_column: int = 0, context: "Context | None" = None) -> torch.Tensor:
    """Counter-addressed N(0,1) block on the device (reference
    random/gaussian.cpp:39-55): entry (i, j) = gaussian_at(seed, stream,
    (first_column + j) * rows + i), returned row-major (rows x cols)."""
    if dtype not in (torch.float64, torch.float32):
        raise ValueError("gaussian blocks are float64 or float32")
    ctx = context or default_context()
    out = torch.empty((rows, cols), dtype=dtype, device=_require_gpu())
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_gaussian_fill(ctx.handle, out.data_ptr(), int(rows), int(cols), int(cols),
                                            _DTYPE_CODE[dtype], int(seed) % _UINT64_MODULUS,
                                            int(stream) % _UINT64_MODULUS, int(first_column) % _UINT64_MODULUS))
    ctx.synchronize()
    return out


def _canonical_uint64_to_int64_transport(value: int, name: str) -> int:
    if isinstance(value, bool) or not isinstance(value, int):
        raise ValueError(f"{name} must be an integer in [0, 2**64 - 1]")
    if value < 0 or value >= _UINT64_MODULUS:
        raise ValueError(f"{name} must be an integer in [0, 2**64 - 1]")
    return value if value < _INT64_SIGN_BIT else value - _UINT64_MODULUS


def _check_inputs(matrix: torch.Tensor, b: torch.Tensor) -> None:
    """reference binding.py:100-110 (device restriction lifted)."""
    if matrix.layout not in (torch.strided, torch.sparse_csr):
        raise ValueError("matrix must use dense strided or sparse CSR storage")
    if b.layout != torch.strided:
        raise ValueError("b must use dense strided storage")
    if matrix.ndim != 2 or b.ndim != 1 or matrix.shape[0] != b.shape[0]:
        raise ValueError("matrix must have shape (m,n) and b must have shape (m,)")
    if matrix.dtype not in (torch.float64, torch.float32, torch.bfloat16) or b.dtype != torch.float64:
        raise ValueError("matrix must be floating point and b must use torch.float64")


def _rhs(b: torch.Tensor, dev: torch.device) -> torch.Tensor:
    return b.to(device=dev, dtype=torch.float64).contiguous()


def _result_tuple(ctx: Context, res_handle: ctypes.c_void_p):
    """Copy an engine result into the reference's 8-tuple (CPU tensors)."""
    lib = ctx._lib
    try:
        cols, rows = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.amfls_result_sizes(res_handle, ctypes.byref(cols), ctypes.byref(rows)))
        sol = np.empty(cols.value, dtype=np.float64)
        met = np.empty(21, dtype=np.float64)
        cnt = np.empty(36, dtype=np.int64)
        tim = np.empty(9, dtype=np.float64)
        tf = np.empty((rows.value, 29), dtype=np.float64)
        ti = np.empty((rows.value, 4), dtype=np.int64)
        ext = np.empty((rows.value, 3), dtype=np.float64)
        st, sr = ctypes.c_int64(), ctypes.c_int64()
        steps = ctypes.c_int32()
        fwd = ctypes.c_double()
        fell = ctypes.c_int32()
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        _lib.check(lib.amfls_result_copy(res_handle, ptr(sol), ptr(met), ptr(cnt), ptr(tim), ptr(tf), ptr(ti),
                                         ctypes.byref(st), ctypes.byref(sr)))
        _lib.check(lib.amfls_result_trace_extension(res_handle, ptr(ext)))
        _lib.check(lib.amfls_result_refinement_steps(res_handle, ctypes.byref(steps)))
        _lib.check(lib.amfls_result_forward_error_estimate(res_handle, ctypes.byref(fwd)))
        _lib.check(lib.amfls_result_fp64_fallback(res_handle, ctypes.byref(fell)))
    finally:
        lib.amfls_result_destroy(res_handle)
    raw = (torch.from_numpy(sol), torch.from_numpy(met), torch.from_numpy(cnt), torch.from_numpy(tim),
           torch.from_numpy(tf), torch.from_numpy(ti), int(st.value), int(sr.value))
    extra = {"refinement_steps": int(steps.value), "forward_error_estimate": float(fwd.value),
             "fp64_fallback": bool(fell.value)}
    return raw, torch.from_numpy(ext), extra


def _decode(raw, ext=None, extra=None) -> SolveResult:
    """reference binding.py:159-255."""
    if not isinstance(raw, tuple) or len(raw) != 8:
        raise RuntimeError("AMFLS binding result must be an eight-item tuple")
    solution, metrics, counters, timings, trace_float, trace_int, status, reason = raw
    if metrics.shape != (21,) or counters.shape != (36,) or timings.shape != (9,):
        raise RuntimeError("AMFLS result tensors have the wrong widths")
    if trace_float.ndim != 2 or trace_float.shape[1] != 29 or trace_int.shape != (trace_float.shape[0], 4):
        raise RuntimeError("AMFLS trace tensors have the wrong widths")
    if trace_int.shape[0] > 0 and not bool(torch.all((trace_int[:, 2] == 0) | (trace_int[:, 2] == 1))):
        raise RuntimeError("AMFLS trace success flags must be zero or one")
    kw = {"solution": solution}
    for i, name in enumerate(_METRIC_NAMES):
        kw[name] = float(metrics[i])
    for i, name in enumerate(_COUNTER_NAMES):
        kw[name] = int(counters[i])
    for i, name in enumerate(_TIMING_NAMES):
        kw[name] = float(timings[i])
    kw.update(trace=trace_float[:, :18], trace_float=trace_float, trace_int=trace_int, status=int(status),
              stop_reason=int(reason), trace_extension=ext)
    kw.update(extra or {})
    return SolveResult(**kw)


def _operator(matrix, value_dtype, context):
    if isinstance(matrix, DeviceMatrix):
        return matrix
    return DeviceMatrix(matrix, value_dtype=value_dtype, context=context)


def run_amfls_raw(matrix, b, regularization=0.0, tolerance=1e-8, failure_probability=1e-6, maximum_epochs=0,
                  maximum_depth=0, maximum_auxiliary_width=0, maximum_basis_size=0, seed=0, stream=0, *,
                  precision="fp64", clock="wall", value_dtype=None, maximum_refinements=8,
                  inner_tolerance=0.0, width_policy=0, context=None, forward_tolerance=0.0,
                  maximum_refinement_iterations=0):
    """The engine call behind torch.ops.amfls.run_amfls; returns (8-tuple, ext, extra)."""
    op = _operator(matrix, value_dtype, context)
    ctx = op.context
    dev = op._val.device
    bd = _rhs(b, dev)
    if bd.numel() != op.rows:
        raise ValueError("right-hand side shape mismatch")
    o = _lib.options(regularization=float(regularization), tolerance=float(tolerance),
                     failure_probability=float(failure_probability), maximum_epochs=int(maximum_epochs),
                     maximum_depth=int(maximum_depth), maximum_auxiliary_width=int(maximum_auxiliary_width),
                     maximum_basis_size=int(maximum_basis_size), seed=int(seed) % _UINT64_MODULUS,
                     stream=int(stream) % _UINT64_MODULUS, precision=_lib.PRECISION[precision],
                     clock_mode=_lib.CLOCK[clock], maximum_refinements=int(maximum_refinements),
                     inner_tolerance=float(inner_tolerance), width_policy=int(width_policy),
                     forward_tolerance=float(forward_tolerance),
                     maximum_refinement_iterations=int(maximum_refinement_iterations))
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve(ctx.handle, op.handle, bd.data_ptr(), 1, ctypes.byref(o), ctypes.byref(res)))
    return _result_tuple(ctx, res)


def run_amfls(matrix, b, *, regularization: float = 0.0, tolerance: float = 1e-8,
              failure_probability: float = 1e-6, maximum_epochs: int = 0, maximum_depth: int = 0,
              maximum_auxiliary_width: int = 0, maximum_basis_size: int = 0, seed: int = 0, stream: int = 0,
              precision: str = "fp64", clock: str = "wall", value_dtype=None, maximum_refinements: int = 8,
              inner_tolerance: float = 0.0, width_policy: int = 0, context: Context | None = None,
              forward_tolerance: float = 0.0, maximum_refinement_iterations: int = 0) -> SolveResult:
    """reference binding.py:259-290 on the HIP engine.

    Engine extensions: precision "mixed" runs an FP32 search and an FP64
    refinement preconditioned by its factorization; forward_tolerance > 0
    keeps refining until the forward-error estimate is at most that value."""
    if not isinstance(matrix, DeviceMatrix):
        _check_inputs(matrix, b)
    seed_t = _canonical_uint64_to_int64_transport(seed, "seed")
    stream_t = _canonical_uint64_to_int64_transport(stream, "stream")
    raw, ext, extra = run_amfls_raw(matrix, b, regularization, tolerance, failure_probability, maximum_epochs,
                                    maximum_depth, maximum_auxiliary_width, maximum_basis_size, seed_t, stream_t,
                                    precision=precision, clock=clock, value_dtype=value_dtype,
                                    maximum_refinements=maximum_refinements, inner_tolerance=inner_tolerance,
                                    width_policy=width_policy, context=context, forward_tolerance=forward_tolerance,
                                    maximum_refinement_iterations=maximum_refinement_iterations)
    return _decode(raw, ext, extra)


def _sweep_options(tolerance, failure_probability, maximum_epochs, maximum_depth, maximum_auxiliary_width,
                   maximum_basis_size, seed, stream, precision, clock, width_policy):
    if precision not in ("fp64", "fp32"):
        raise ValueError("regularization sweeps run in fp64 or fp32 precision")
    return _lib.options(tolerance=float(tolerance), failure_probability=float(failure_probability),
                        maximum_epochs=int(maximum_epochs), maximum_depth=int(maximum_depth),
                        maximum_auxiliary_width=int(maximum_auxiliary_width),
                        maximum_basis_size=int(maximum_basis_size), seed=int(seed) % _UINT64_MODULUS,
                        stream=int(stream) % _UINT64_MODULUS, precision=_lib.PRECISION[precision],
                        clock_mode=_lib.CLOCK[clock], width_policy=int(width_policy))


def run_amfls_sweep(matrix, b, lambdas, *, tolerance: float = 1e-8, failure_probability: float = 1e-6,
                    maximum_epochs: int = 0, maximum_depth: int = 0, maximum_auxiliary_width: int = 0,
                    maximum_basis_size: int = 0, seed: int = 0, stream: int = 0, precision: str = "fp64",
                    clock: str = "wall", value_dtype=None, width_policy: int = 0,
                    context: Context | None = None) -> list[SolveResult]:
    """One certified solve per regularization in `lambdas`, all served by one
    block Golub-Kahan process (engine extension; the reference runs
    run_amfls once per lambda, binding.py:259-290)."""
    op = _operator(matrix, value_dtype, context)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    if bd.numel() != op.rows:
        raise ValueError("right-hand side shape mismatch")
    lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64).ravel())
    if lam.size == 0:
        raise ValueError("at least one regularization is required")
    o = _sweep_options(tolerance, failure_probability, maximum_epochs, maximum_depth, maximum_auxiliary_width,
                       maximum_basis_size, seed, stream, precision, clock, width_policy)
    res = (ctypes.c_void_p * lam.size)()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_sweep(ctx.handle, op.handle, bd.data_ptr(), 1,
                                          lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), lam.size,
                                          ctypes.byref(o), res))
    handles = [ctypes.c_void_p(res[j]) for j in range(lam.size)]
    out = []
    for j, h in enumerate(handles):
        try:
            out.append(_decode(*_result_tuple(ctx, h)))
        except Exception:
            for k in handles[j + 1:]:
                ctx._lib.amfls_result_destroy(k)
            raise
    return out


def run_amfls_batch(matrices, rhs, lambdas, *, workers: int = 6, tolerance: float = 1e-8,
                    failure_probability: float = 1e-6, maximum_epochs: int = 0, maximum_depth: int = 0,
                    maximum_auxiliary_width: int = 0, maximum_basis_size: int = 0, seed: int = 0, stream: int = 0,
                    precision: str = "fp64", clock: str = "wall", width_policy: int = 0,
                    context: Context | None = None) -> list[list[SolveResult]]:
    """Independent CSR problems (DeviceMatrix list, device right-hand sides,
    lambdas of shape problems x L) solved concurrently by `workers` host
    threads with a stream each; every problem runs a regularization sweep
    (engine extension)."""
    mats = list(matrices)
    bs = [torch.as_tensor(b) for b in rhs]
    lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64))
    if lam.ndim != 2 or lam.shape[0] != len(mats) or len(bs) != len(mats) or lam.shape[1] == 0:
        raise ValueError("lambdas must have shape (problems, L) with L > 0, one right-hand side per problem")
    ctx = context or (mats[0].context if mats else default_context())
    keep = []
    for M, b in zip(mats, bs):
        if not isinstance(M, DeviceMatrix) or M.kind != "csr":
            raise ValueError("batched solves take CSR DeviceMatrix operators")
        if b.numel() != M.rows:
            raise ValueError("right-hand side shape mismatch")
        keep.append(_rhs(b, M._val.device))
    o = _sweep_options(tolerance, failure_probability, maximum_epochs, maximum_depth, maximum_auxiliary_width,
                       maximum_basis_size, seed, stream, precision, clock, width_policy)
    P, L = lam.shape
    ops = (ctypes.c_void_p * max(P, 1))(*[M.handle.value for M in mats])
    bp = (ctypes.c_void_p * max(P, 1))(*[b.data_ptr() for b in keep])
    res = (ctypes.c_void_p * max(P * L, 1))()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_batch(ctx.handle, P, ops, bp, lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          L, ctypes.byref(o), int(workers), res))
    handles = [ctypes.c_void_p(res[k]) for k in range(P * L)]
    out, flat = [], []
    try:
        for h in handles:
            taken = ctypes.c_void_p(h.value)
            h.value = None  # _result_tuple releases it, also when it raises
            flat.append(_decode(*_result_tuple(ctx, taken)))
    finally:
        for h in handles:
            if h.value:
                ctx._lib.amfls_result_destroy(h)
    for i in range(P):
        out.append(flat[i * L:(i + 1) * L])
    return out


def run_fixed_rbgk(matrix, b, *, auxiliary_width: int, maximum_depth: int, maximum_basis_size: int = 0,
                   regularization: float = 0.0, tolerance: float = 1e-8, failure_probability: float = 1e-6,
                   seed: int = 0, stream: int = 0, clock: str = "wall", value_dtype=None,
                   context: Context | None = None) -> SolveResult:
    """reference binding.py run_fixed_rbgk / fixed_rbgk.cpp:18-209 on the HIP engine."""
    if not isinstance(matrix, DeviceMatrix):
        _check_inputs(matrix, b)
    op = _operator(matrix, value_dtype, context)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_fixed_rbgk(ctx.handle, op.handle, bd.data_ptr(), 1, float(regularization),
                                               float(tolerance), float(failure_probability), int(auxiliary_width),
                                               int(maximum_depth), int(maximum_basis_size),
                                               int(seed) % _UINT64_MODULUS, int(stream) % _UINT64_MODULUS,
                                               _lib.CLOCK[clock], ctypes.byref(res)))
    raw, ext, extra = _result_tuple(ctx, res)
    return _decode(raw, ext, extra)


def validate_candidate(matrix, b, x, *, regularization: float = 0.0, tolerance: float = 1e-8, value_dtype=None,
                       context: Context | None = None) -> SolveResult:
    """Certify a given solution x of min ||Ax - b||^2 + lambda ||x||^2 with
    the engine's FP64 validator (reference candidate_validation.cpp:1176-1378
    validate_original_candidate): the result carries x, its residual and
    gradient norms, the certified backward-error / energy bounds and the
    status the certificate implies at `tolerance`."""
    if not isinstance(matrix, DeviceMatrix):
        _check_inputs(matrix, b)
    op = _operator(matrix, value_dtype, context)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    xd = _rhs(x, op._val.device)
    if xd.numel() != op.cols:
        raise ValueError(f"the candidate has {xd.numel()} entries, the operator {op.cols} columns")
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_validate_candidate(ctx.handle, op.handle, bd.data_ptr(), 1, xd.data_ptr(), 1,
                                                 float(regularization), float(tolerance), ctypes.byref(res)))
    raw, ext, extra = _result_tuple(ctx, res)
    return _decode(raw, ext, extra)


def run_lsqr(matrix, b, *, regularization: float = 0.0, tolerance: float = 1e-8, maximum_iterations: int = 0,
             clock: str = "wall", value_dtype=None, context: Context | None = None) -> SolveResult:
    """reference binding.py run_lsqr / lsqr.cpp:84-279 on the HIP engine."""
    if not isinstance(matrix, DeviceMatrix):
        _check_inputs(matrix, b)
    op = _operator(matrix, value_dtype, context)
    ctx = op.context
    bd = _rhs(b, op._val.device)
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve_lsqr(ctx.handle, op.handle, bd.data_ptr(), 1, float(regularization),
                                         float(tolerance), int(maximum_iterations), _lib.CLOCK[clock],
                                         ctypes.byref(res)))
    raw, ext, extra = _result_tuple(ctx, res)
    return _decode(raw, ext, extra)
