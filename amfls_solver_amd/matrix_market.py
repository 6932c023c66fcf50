"""Matrix Market files as CSR operators and right-hand sides.

The reference's application study feeds sparse problems from Matrix Market
files (experiments/runners/run_application_study.py:80-160).  This module is
a self-contained reader for the subset such problems use:

* banner ``%%MatrixMarket matrix <coordinate|array> <real|integer|pattern>
  <general|symmetric|skew-symmetric>`` (case-insensitive); complex and
  Hermitian fields are refused;
* optional gzip compression (a ``.gz`` suffix);
* 1-based coordinate triplets, or column-major dense arrays (only the lower
  triangle is stored for symmetric arrays).

``MatrixMarketHeader`` describes a file, ``load_sparse(path)`` returns a
canonical float64 CSR matrix (repeated coordinates added, stored zeros
dropped, column indices increasing within each row), ``load_vector(path,
length, unit=True)`` a right-hand side (scaled to unit 2-norm by default,
the scaling the study applies), and ``load_csr(path)`` the engine's device
operator.  Malformed input raises ``MatrixMarketError`` (a ValueError); a
missing file raises FileNotFoundError.
"""
from __future__ import annotations

import gzip
import io
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import scipy.sparse

_FORMATS = ("coordinate", "array")
_FIELDS = ("real", "integer", "pattern")
_SYMMETRIES = ("general", "symmetric", "skew-symmetric")


class MatrixMarketError(ValueError):
    """The file is not a Matrix Market matrix this reader accepts."""


@dataclass(frozen=True)
class MatrixMarketHeader:
    layout: str      # "coordinate" or "array"
    field: str       # "real", "integer" or "pattern"
    symmetry: str    # "general", "symmetric" or "skew-symmetric"
    rows: int
    cols: int
    entries: int     # stored entries (coordinate) or stored values (array)


def _text(path) -> tuple[Path, io.TextIOBase]:
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(path)
    raw = gzip.open(path, "rb") if path.suffix == ".gz" else open(path, "rb")
    return path, io.TextIOWrapper(raw, encoding="ascii", errors="strict")


def _parse_banner(line: str, path: Path) -> tuple[str, str, str]:
    words = line.strip().lower().split()
    if len(words) != 5 or words[0] != "%%matrixmarket" or words[1] != "matrix":
        raise MatrixMarketError(f"{path}: first line is not a Matrix Market matrix banner")
    layout, field, symmetry = words[2:]
    if field in ("complex", "hermitian") or symmetry == "hermitian":
        raise MatrixMarketError(f"{path}: complex data cannot be solved in real arithmetic")
    if layout not in _FORMATS or field not in _FIELDS or symmetry not in _SYMMETRIES:
        raise MatrixMarketError(f"{path}: unsupported banner {' '.join(words[2:])!r}")
    if layout == "array" and field == "pattern":
        raise MatrixMarketError(f"{path}: a dense array cannot have a pattern field")
    return layout, field, symmetry


def _read(path) -> tuple[MatrixMarketHeader, np.ndarray]:
    """Header and the data section as a float64 array (one row per line)."""
    path, f = _text(path)
    with f:
        layout, field, symmetry = _parse_banner(f.readline(), path)
        size_line = ""
        for line in f:
            s = line.strip()
            if s and not s.startswith("%"):
                size_line = s
                break
        try:
            dims = [int(w) for w in size_line.split()]
        except ValueError:
            dims = []
        want = 3 if layout == "coordinate" else 2
        if len(dims) != want or min(dims) < 0:
            raise MatrixMarketError(f"{path}: size line {size_line!r} does not give {want} counts")
        rows, cols = dims[0], dims[1]
        if layout == "coordinate":
            entries = dims[2]
        elif symmetry == "general":
            entries = rows * cols
        else:
            if rows != cols:
                raise MatrixMarketError(f"{path}: a {symmetry} array must be square")
            entries = rows * (rows + 1) // 2 - (rows if symmetry == "skew-symmetric" else 0)
        width = (2 if field == "pattern" else 3) if layout == "coordinate" else 1
        body = f.read()
    data = np.loadtxt(io.StringIO(body), comments="%", dtype=np.float64, ndmin=2) if entries else \
        np.zeros((0, width))
    if data.shape != (entries, width) and not (entries == 0 and data.size == 0):
        raise MatrixMarketError(f"{path}: expected {entries} data lines of {width} numbers, "
                                f"found an array of shape {data.shape}")
    return MatrixMarketHeader(layout, field, symmetry, rows, cols, entries), data.reshape(entries, width)


def _coordinates(h: MatrixMarketHeader, data: np.ndarray, path) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """0-based (row, col, value) triplets with the symmetric half expanded."""
    if h.layout == "coordinate":
        idx = data[:, :2]
        if not np.all(idx == np.round(idx)):
            raise MatrixMarketError(f"{path}: coordinates must be integers")
        r = idx[:, 0].astype(np.int64) - 1
        c = idx[:, 1].astype(np.int64) - 1
        v = np.ones(len(r)) if h.field == "pattern" else data[:, 2].copy()
        if len(r) and (r.min() < 0 or c.min() < 0 or r.max() >= h.rows or c.max() >= h.cols):
            raise MatrixMarketError(f"{path}: a coordinate lies outside the {h.rows} x {h.cols} matrix")
    else:
        vals = data[:, 0]
        if h.symmetry == "general":
            c, r = np.divmod(np.arange(h.rows * h.cols, dtype=np.int64), h.rows)  # column-major order
        else:
            first = 1 if h.symmetry == "skew-symmetric" else 0
            cc, rr = np.meshgrid(np.arange(h.cols), np.arange(h.rows), indexing="ij")
            keep = rr >= cc + first                                            # lower triangle, by column
            r, c = rr[keep].astype(np.int64), cc[keep].astype(np.int64)
        v = vals.copy()
    if h.symmetry != "general":
        if h.rows != h.cols:
            raise MatrixMarketError(f"{path}: a {h.symmetry} matrix must be square")
        off = r != c
        if h.symmetry == "skew-symmetric" and np.any(~off):
            raise MatrixMarketError(f"{path}: a skew-symmetric matrix stores no diagonal")
        sign = -1.0 if h.symmetry == "skew-symmetric" else 1.0
        r, c, v = (np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, sign * v[off]]))
    return r, c, v


def load_sparse(path) -> scipy.sparse.csr_matrix:
    """Canonical float64 CSR matrix of a Matrix Market file."""
    h, data = _read(path)
    if h.rows == 0 or h.cols == 0:
        raise MatrixMarketError(f"{path}: the matrix is empty ({h.rows} x {h.cols})")
    r, c, v = _coordinates(h, data, path)
    if not np.isfinite(v).all():
        raise MatrixMarketError(f"{path}: the matrix holds a NaN or an infinity")
    A = scipy.sparse.csr_matrix((v, (r, c)), shape=(h.rows, h.cols), dtype=np.float64)
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    if A.nnz == 0:
        raise MatrixMarketError(f"{path}: every stored entry is zero")
    return A


def load_vector(path, length: int, *, unit: bool = True) -> np.ndarray:
    """A right-hand side of `length` entries from a one-column (or one-row)
    Matrix Market file, scaled to unit 2-norm unless unit=False."""
    h, data = _read(path)
    if sorted((h.rows, h.cols)) != [1, length]:
        raise MatrixMarketError(f"{path}: holds a {h.rows} x {h.cols} matrix, not a vector of {length} entries")
    r, c, v = _coordinates(h, data, path)
    b = np.zeros(length)
    np.add.at(b, r if h.cols == 1 else c, v)
    if not np.isfinite(b).all():
        raise MatrixMarketError(f"{path}: the vector holds a NaN or an infinity")
    norm = float(np.linalg.norm(b))
    if norm == 0.0:
        raise MatrixMarketError(f"{path}: the vector is zero, so the problem has the trivial solution")
    return np.ascontiguousarray(b / norm if unit else b)


def load_csr(path, *, value_dtype=None, context=None):
    """Device CSR operator of a Matrix Market file (float64 values unless
    value_dtype says otherwise)."""
    import torch

    from .binding import DeviceMatrix
    A = load_sparse(path)
    t = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                torch.from_numpy(A.indices.astype(np.int64)),
                                torch.from_numpy(A.data), size=A.shape)
    return DeviceMatrix(t, value_dtype=value_dtype or torch.float64, context=context)
