"""Matrix Market ingestion for CSR problems (the reference's application
study loads its sparse inputs this way, experiments/runners/
run_application_study.py:80-160).

read_matrix_market(path) returns a float64 scipy CSR matrix with duplicates
summed, explicit zeros removed and column indices sorted; read_rhs(path,
rows) returns the right-hand side scaled to unit norm.  Both accept
gzip-compressed files (".gz") and refuse complex, empty or nonfinite data
with ValueError, and missing files with FileNotFoundError.  load_csr(path)
builds the engine's device operator from such a file.
"""
from __future__ import annotations

import gzip
from pathlib import Path

import numpy as np
import scipy.io
import scipy.sparse


def _mmread(path: Path):
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(path)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rb") as handle:
        return scipy.io.mmread(handle, spmatrix=True)


def read_matrix_market(path) -> scipy.sparse.csr_matrix:
    value = _mmread(path)
    if scipy.sparse.issparse(value):
        if np.iscomplexobj(value.data):
            raise ValueError(f"complex matrices are not supported: {path}")
        A = value.astype(np.float64).tocsr()
    else:
        dense = np.asarray(value)
        if np.iscomplexobj(dense):
            raise ValueError(f"complex matrices are not supported: {path}")
        A = scipy.sparse.csr_matrix(dense.astype(np.float64))
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    if A.ndim != 2 or min(A.shape) <= 0:
        raise ValueError(f"matrix must be two-dimensional and nonempty: {path}")
    if A.nnz == 0 or not np.isfinite(A.data).all():
        raise ValueError(f"matrix must hold finite nonzero entries: {path}")
    return A


def read_rhs(path, rows: int) -> np.ndarray:
    value = _mmread(path)
    arr = value.toarray() if scipy.sparse.issparse(value) else np.asarray(value)
    if np.iscomplexobj(arr):
        raise ValueError(f"complex right-hand sides are not supported: {path}")
    if arr.shape not in {(rows,), (rows, 1), (1, rows)}:
        raise ValueError(f"right-hand side must be a vector of length {rows}: {path}")
    b = np.asarray(arr, dtype=np.float64).reshape(-1)
    if not np.isfinite(b).all():
        raise ValueError(f"right-hand side must be finite: {path}")
    nrm = float(np.linalg.norm(b))
    if nrm == 0.0:
        raise ValueError(f"right-hand side cannot be zero: {path}")
    return np.ascontiguousarray(b / nrm)


def load_csr(path, *, value_dtype=None, context=None):
    """Device CSR operator of a Matrix Market file (float64 values unless
    value_dtype says otherwise)."""
    import torch

    from .binding import DeviceMatrix
    A = read_matrix_market(path)
    t = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                torch.from_numpy(A.indices.astype(np.int64)),
                                torch.from_numpy(A.data), size=A.shape)
    return DeviceMatrix(t, value_dtype=value_dtype or torch.float64, context=context)
