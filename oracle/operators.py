"""Built-in FP64 operators (oracle).

Follows src/algorithms/mathematics/operators/:
* ``CsrOperator``       -- sparse_csr_operator.cpp:14-269
* ``DenseOperator``     -- dense_operator.cpp:20-186 (column-major storage)
* ``RowMajorDenseOperator`` -- row_major_dense_operator.cpp:23-257
Every constructor scans |A| into outward-rounded ||A|| bounds
(upper = min(sqrt(||A||_1 ||A||_inf), ||A||_F), lower = max(max|a_ij|,
max row/column 2-norm)) and the validation model's rounding-step counts.

Blocks are numpy arrays of shape (rows, s); storage order is immaterial to
numpy.  Sums use numpy reductions instead of the reference's sequential loops;
each still obeys the gamma_k bound that positive_sum_{upper,lower}_bound
assume, so the bounds stay valid and agree with the reference to the last
few ulps.

Pinned by tests/test_oracle_solvers.py: CSR products against the dense
product of the same matrix, the norm bounds against numpy's 2-norm
(lower <= ||A||_2 <= upper, and upper >= || |A| ||_2), the step counts
2 max_row_nnz + 2 / 2 max_col_nnz + 2, the rank-one closed form
||u v^T|| = ||u|| ||v||, the CSR structure checks (duplicates, range,
nonfinite values) and the long-double refinement product exact on small
integers (test_operator_products_and_norm_bounds,
test_csr_constructor_rejects_bad_structure,
test_long_double_refinement_is_exact_on_small_integers).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from . import fpbounds as fp

INF = math.inf
LD = np.longdouble


class MatrixOperator:
    """matrix_operator.hpp:29-89 default behaviour (no validation model)."""

    rows: int
    cols: int

    def apply(self, x):  # pragma: no cover - abstract
        raise NotImplementedError

    def apply_transpose(self, y):  # pragma: no cover - abstract
        raise NotImplementedError

    def relative_block_product_cost(self, s: int) -> float:
        return float(s) if s > 0 else INF

    def validation_error_model(self):
        return (INF, 0.0, 0, 0)

    def supports_validation_refinement(self) -> bool:
        return False

    def apply_validation_refinement(self, x):
        return None

    def apply_transpose_validation_refinement(self, y):
        return None


def _max_bound_by_group(values: np.ndarray, ops: np.ndarray, fn) -> float:
    """max_i fn(values[i], ops[i]) for fn monotone increasing in value.

    The reference takes this maximum row by row; grouping by operation count
    gives the same maximum with one bound evaluation per distinct count.
    """
    if values.size == 0:
        return 0.0
    best = 0.0
    for k in np.unique(ops):
        sel = ops == k
        best = max(best, fn(float(values[sel].max()), int(k)))
    return best


def _norm_bounds(max_entry, row_sums, row_ops, row_sq, col_sums, col_ops, col_sq):
    """Shared tail of the three constructors (e.g. sparse_csr_operator.cpp:90-129)."""
    max_row_sum_upper = _max_bound_by_group(row_sums, row_ops, fp.positive_sum_upper_bound)
    max_col_sum_upper = _max_bound_by_group(col_sums, col_ops, fp.positive_sum_upper_bound)
    two_norm_lower = 0.0
    for sq, ops in ((row_sq, row_ops), (col_sq, col_ops)):
        lows = fp.positive_sum_lower_bound_array(sq, ops)
        if lows.size:
            two_norm_lower = max(two_norm_lower, float(fp.downward_sqrt_array(lows).max()))
    fro_terms = fp.positive_fma_square_sum_upper_bound_array(col_sq, col_ops)
    fro_sq_upper = 0.0
    for t in fro_terms:
        fro_sq_upper = fp.upward_add(fro_sq_upper, float(t))
    one_inf = fp.upward_multiply(fp.upward_sqrt(max_col_sum_upper), fp.upward_sqrt(max_row_sum_upper))
    upper = min(one_inf, fp.upward_sqrt(fro_sq_upper))
    lower = max(max_entry, two_norm_lower)
    if not math.isfinite(upper) or lower > upper:
        return INF, 0.0
    return upper, lower


class CsrOperator(MatrixOperator):
    def __init__(self, rows, cols, row_offsets, column_indices, values):
        if rows <= 0 or cols <= 0:
            raise ValueError("CSR operator dimensions are invalid")
        rp = np.asarray(row_offsets, dtype=np.int64)
        ci = np.asarray(column_indices, dtype=np.int64)
        va = np.asarray(values, dtype=np.float64)
        nnz = int(va.size)
        if rp.size != rows + 1 or rp[0] != 0 or rp[-1] != nnz or ci.size != nnz:
            raise ValueError("CSR row offsets do not span the values")
        if np.any(np.diff(rp) < 0):
            raise ValueError("CSR row offsets must be nondecreasing")
        if nnz and (ci.min() < 0 or ci.max() >= cols):
            raise ValueError("CSR column index is out of range")
        if not np.all(np.isfinite(va)):
            raise ValueError("CSR values must be finite")
        row_of = np.repeat(np.arange(rows, dtype=np.int64), np.diff(rp))
        if nnz:
            key = row_of * cols + ci
            ks = np.sort(key)
            if np.any(ks[1:] == ks[:-1]):
                raise ValueError("CSR rows must not contain duplicate columns")
        self.rows, self.cols = rows, cols
        self.row_offsets, self.column_indices, self.values = rp, ci, va
        self.nnz = nnz
        self.mat = sp.csr_matrix((va, ci, rp), shape=(rows, cols))
        self.mat_t = self.mat.T.tocsr()
        mag = np.abs(va)
        row_nnz = np.diff(rp)
        col_nnz = np.bincount(ci, minlength=cols).astype(np.int64)
        self.max_row_nnz = int(row_nnz.max()) if rows else 0
        self.max_col_nnz = int(col_nnz.max()) if cols else 0
        row_sums = np.bincount(row_of, weights=mag, minlength=rows)
        row_sq = np.bincount(row_of, weights=mag * mag, minlength=rows)
        col_sums = np.bincount(ci, weights=mag, minlength=cols)
        col_sq = np.bincount(ci, weights=mag * mag, minlength=cols)
        max_entry = float(mag.max()) if nnz else 0.0
        self.norm_upper, self.norm_lower = _norm_bounds(
            max_entry, row_sums, row_nnz, row_sq, col_sums, col_nnz, col_sq)

    def apply(self, x):
        return self.mat @ np.asarray(x, dtype=np.float64)

    def apply_transpose(self, y):
        return self.mat_t @ np.asarray(y, dtype=np.float64)

    def validation_error_model(self):
        return (self.norm_upper, self.norm_lower, 2 * self.max_row_nnz + 2,
                2 * self.max_col_nnz + 2)

    def supports_validation_refinement(self):
        return True

    def _ld_spmv(self, rp, ci, va, x):
        prod = va.astype(LD) * np.asarray(x, dtype=LD)[ci]
        cs = np.concatenate([[LD(0)], np.cumsum(prod, dtype=LD)]) if prod.size else np.zeros(1, LD)
        # row sums as differences of a long-double prefix sum would add error;
        # use an explicit segmented reduction instead.
        out = np.zeros(rp.size - 1, dtype=LD)
        nz = np.diff(rp) > 0
        if prod.size:
            out[nz] = np.add.reduceat(prod, rp[:-1][nz])
        del cs
        return out

    def apply_validation_refinement(self, x):
        y = self._ld_spmv(self.row_offsets, self.column_indices, self.values, x)
        return y if np.all(np.isfinite(y)) else None

    def apply_transpose_validation_refinement(self, y):
        t = self.mat_t
        x = self._ld_spmv(t.indptr.astype(np.int64), t.indices.astype(np.int64), t.data, y)
        return x if np.all(np.isfinite(x)) else None


class _DenseBase(MatrixOperator):
    def __init__(self, a: np.ndarray):
        a = np.asarray(a, dtype=np.float64)
        m, n = a.shape
        if m <= 0 or n <= 0:
            raise ValueError("dense operator input is invalid")
        self.rows, self.cols = m, n
        self.a = a
        mag = np.abs(a)
        certifiable = bool(np.all(np.isfinite(mag)))
        mag = np.where(np.isfinite(mag), mag, 0.0)
        row_sums, row_sq = mag.sum(axis=1), (mag * mag).sum(axis=1)
        col_sums, col_sq = mag.sum(axis=0), (mag * mag).sum(axis=0)
        up, lo = _norm_bounds(float(mag.max()), row_sums, np.full(m, n, np.int64), row_sq,
                              col_sums, np.full(n, m, np.int64), col_sq)
        self.norm_upper, self.norm_lower = (up, lo) if certifiable else (INF, 0.0)

    def apply(self, x):
        return self.a @ np.asarray(x, dtype=np.float64)

    def apply_transpose(self, y):
        return self.a.T @ np.asarray(y, dtype=np.float64)

    def relative_block_product_cost(self, s: int) -> float:
        return 1.0 if s > 0 else INF

    def validation_error_model(self):
        return (self.norm_upper, self.norm_lower, 2 * self.cols + 2, 2 * self.rows + 2)

    def supports_validation_refinement(self):
        return True

    def apply_validation_refinement(self, x):
        y = self.a.astype(LD) @ np.asarray(x, dtype=LD)
        return y if np.all(np.isfinite(y)) else None

    def apply_transpose_validation_refinement(self, y):
        x = self.a.T.astype(LD) @ np.asarray(y, dtype=LD)
        return x if np.all(np.isfinite(x)) else None


class DenseOperator(_DenseBase):
    """Column-major dense A (dense_operator.cpp)."""


class RowMajorDenseOperator(_DenseBase):
    """Row-major dense A (row_major_dense_operator.cpp); same mathematics."""


class DiagonalOperator(MatrixOperator):
    """The README's matrix-free example (README.md:534-579)."""

    def __init__(self, diagonal):
        self.d = np.asarray(diagonal, dtype=np.float64)
        self.rows = self.cols = int(self.d.size)
        self.norm = float(np.max(np.abs(self.d)))

    def apply(self, x):
        return self.d[:, None] * np.asarray(x).reshape(self.rows, -1)

    def apply_transpose(self, y):
        return self.apply(y)

    def validation_error_model(self):
        return (self.norm, self.norm, 1, 1)
