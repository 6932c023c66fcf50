"""Seeded synthetic inputs for the AMFLS benchmark configurations.

* ``dense_gd``   -- dense A = G D, G i.i.d. N(0,1), D log-spaced (the
                    2000 x 500 and 1e6 x 2000 dense configurations)
* ``csr_rows``   -- CSR with a fixed number of distinct uniform columns per
                    row, N(0,1) values and log-uniform row scales (the
                    200000 x 20000 configuration and its scaled variants)
* ``csr_batch_problem`` -- one problem of the batched configuration

Small and medium problems are generated on the host with numpy's PCG64
(seeds recorded by the callers); the 1e6 x 2000 dense matrix is generated on
the device with the counter-addressed Gaussian stream of the engine
(amfls_gaussian_fill) so that it never exists in host memory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CsrProblem:
    rows: int
    cols: int
    row_ptr: np.ndarray  # int64, rows + 1
    col_idx: np.ndarray  # int32, nnz
    values: np.ndarray   # float32 (stored values define A)
    b: np.ndarray        # float64
    regularization: float

    @property
    def nnz(self) -> int:
        return int(self.values.size)

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values.astype(np.float64), self.col_idx, self.row_ptr),
                             shape=(self.rows, self.cols))


def distinct_columns(rng: np.random.Generator, rows: int, cols: int, per_row: np.ndarray) -> np.ndarray:
    """Uniform distinct column indices per row (sorted within each row)."""
    per_row = np.asarray(per_row, dtype=np.int64)
    kmax = int(per_row.max()) if per_row.size else 0
    idx = rng.integers(0, cols, size=(rows, kmax), dtype=np.int64)
    mask = np.arange(kmax)[None, :] < per_row[:, None]
    while True:
        s = np.sort(np.where(mask, idx, cols + np.arange(kmax)[None, :]), axis=1)
        dup = (s[:, 1:] == s[:, :-1]) & (s[:, 1:] < cols)
        bad = np.nonzero(dup.any(axis=1))[0]
        if bad.size == 0:
            idx = s
            break
        idx[bad] = rng.integers(0, cols, size=(bad.size, kmax), dtype=np.int64)
    return idx[mask].astype(np.int32)  # row-major order, sorted per row


def csr_rows(rows: int, cols: int, per_row, *, seed: int = 1, scale_range=(1e-3, 1.0),
             regularization: float = 1e-3, rhs: str = "gaussian") -> CsrProblem:
    """CSR with `per_row` distinct uniform columns, N(0,1) values, rows scaled by
    log-uniform factors in `scale_range`; b ~ N(0,1).  Values are stored as
    float32 and those stored values are the matrix A."""
    rng = np.random.default_rng(seed)
    per = np.full(rows, per_row, dtype=np.int64) if np.isscalar(per_row) else np.asarray(per_row, np.int64)
    ci = distinct_columns(rng, rows, cols, per)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(per, out=rp[1:])
    vals = rng.standard_normal(int(rp[-1]))
    lo, hi = np.log10(scale_range[0]), np.log10(scale_range[1])
    scale = 10.0 ** rng.uniform(lo, hi, size=rows)
    vals *= np.repeat(scale, per)
    b = rng.standard_normal(rows)
    return CsrProblem(rows, cols, rp, ci, vals.astype(np.float32), b, regularization)


def config1(seed: int = 1) -> CsrProblem:
    """200000 x 20000 CSR, 12 distinct columns per row, lambda = 1e-3."""
    return csr_rows(200_000, 20_000, 12, seed=seed, regularization=1e-3)


def csr_batch_problem(index: int, seed: int = 3) -> CsrProblem:
    """Problem `index` of the batched configuration: m ~ U[5000, 200000],
    n ~ U[500, min(10000, m)], nnz per row ~ U[4, 32]."""
    rng = np.random.default_rng([seed, index])
    m = int(rng.integers(5000, 200_001))
    n = int(rng.integers(500, min(10_000, m) + 1))
    per = rng.integers(4, 33, size=m)
    per = np.minimum(per, n)
    p = csr_rows(m, n, per, seed=int(rng.integers(0, 2**31)), scale_range=(1.0, 1.0), regularization=0.0)
    return p


def dense_gd(rows: int, cols: int, decades: float, *, seed: int = 0, noise: float = 1e-6,
             relative_noise: bool = False):
    """A = G D (column-major float64), x* ~ N(0,1), b = A x* + noise * e."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((rows, cols))
    d = np.logspace(0.0, -decades, cols)
    a = np.asfortranarray(g * d[None, :])
    xs = rng.standard_normal(cols)
    ax = a @ xs
    e = rng.standard_normal(rows)
    if relative_noise:
        b = ax + noise * np.linalg.norm(ax) * e / np.linalg.norm(e)
    else:
        b = ax + noise * e
    return a, b, xs


def config0(seed: int = 0):
    """2000 x 500 dense, D = logspace(0, -4), b = A x* + 1e-6 N(0,1)."""
    return dense_gd(2000, 500, 4.0, seed=seed, noise=1e-6)


def config2_device(rows: int = 1_000_000, cols: int = 2000, *, storage=None, decades: float = 6.0, seed: int = 0,
                   noise: float = 1e-8):
    """Dense tall A = G D built on the GPU: G_ij = gaussian_at(seed, 0, j*rows + i)
    (the counter-addressed stream of reference random/gaussian.cpp:39-55), D
    log-spaced from 1 to 10^-decades, stored as `storage` (bfloat16 by
    default; the stored values are A); x* = gaussian_at(seed, 1, .),
    e = gaussian_at(seed, 2, .), b = A x* + noise ||A x*|| e / ||e|| with A x*
    computed in FP64 from the stored A.  Returns (A, b, x*) as CUDA tensors,
    A row-major (rows x cols)."""
    import torch

    from .binding import DeviceMatrix, gaussian_block
    storage = storage or torch.bfloat16
    d = torch.logspace(0.0, -decades, cols, dtype=torch.float64, device="cuda")
    a = torch.empty((rows, cols), dtype=storage, device="cuda")
    chunk = max(1, min(cols, (1 << 26) // max(rows, 1)))  # columns per generated slab (<= 512 MB)
    for j0 in range(0, cols, chunk):
        j1 = min(cols, j0 + chunk)
        g = gaussian_block(rows, j1 - j0, seed, 0, first_column=j0)  # entry (i, j) = counter (j0 + j) rows + i
        a[:, j0:j1] = (g * d[None, j0:j1]).to(storage)
        del g
    xs = gaussian_block(cols, 1, seed, 1).reshape(-1)
    e = gaussian_block(rows, 1, seed, 2).reshape(-1)
    ax = DeviceMatrix(a, value_dtype=storage).apply(xs).reshape(-1)
    b = ax + noise * torch.linalg.vector_norm(ax) * e / torch.linalg.vector_norm(e)
    return a, b, xs
