"""Writes tests/golden/c2_solution.npz: the FP64 least-squares solution of the
1,000,000 x 2000 dense workload, for the forward-error checks of the GPU
solves (run from the repository root: python tests/golden/make_c2_fixture.py;
about five minutes on 8 cores and 12 GB of memory).

The workload, exactly as amfls_solver_amd.problems.config2_device builds it
on the GPU:
  G_ij = gaussian_at(seed 0, stream 0, j * rows + i)   (counter-addressed N(0,1))
  D_jj = 10^(-6 j / 1999)                              (log-spaced, 1 .. 1e-6)
  A    = bfloat16(G D)  -- rounded double -> float -> bfloat16 (round to
                           nearest even each time, as a torch conversion does);
                           the stored bfloat16 values ARE the matrix A
  x*   = gaussian_at(0, 1, .),  e = gaussian_at(0, 2, .)
  b    = A x* + 1e-8 ||A x*|| e / ||e||,  lambda = 0.

The oracle x_LS = argmin ||A x - b|| is computed in FP64 by a Householder
tall-skinny QR (TSQR) of [A | A x* | e] over row blocks of the stored A upcast
exactly to FP64: the last two columns of R give Q^T (A x*) and Q^T e, so
x_LS = R_11^{-1} (Q^T b)_{1:n}.  Pinned by what the mathematics fixes:
  * the FP64 normal-equations residual ||A^T (b - A x_LS)|| / (||A||_F ||r||),
    recomputed in a second pass over A, must be at FP64 rounding level;
  * one corrected-semi-normal-equations step delta = R^{-1} R^{-T} A^T r must
    change x_LS by a relative amount far below 1e-8 (its size is recorded);
  * ||r|| computed from the explicit residual must equal |R_{n+1,n+1}| of the
    augmented QR (the LS residual norm) to working accuracy.
A bit-level checksum of the stored A is recorded so the GPU test can show it
solved the same matrix (one flipped bfloat16 entry changes x_LS far more than
1e-8).
"""
from __future__ import annotations

import mmap
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

ROWS, COLS, DECADES, SEED, NOISE = 1_000_000, 2000, 6.0, 0, 1e-8
BLOCK = 50_000          # rows per TSQR block
SUB = 5_000             # rows generated at once inside a worker
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_solution.npz")

_shared = None  # float32 view of the stored A (ROWS x COLS, row-major), set before the workers fork


def _scales():
    import torch
    return torch.logspace(0.0, -DECADES, COLS, dtype=torch.float64).numpy()


def _block(i0: int):
    """Generate rows [i0, i0 + BLOCK) of the stored A into the shared array;
    returns the three bit checksums of the block."""
    import torch
    from oracle import rng
    torch.set_num_threads(1)
    d = _scales()
    cols = np.arange(COLS, dtype=np.uint64)
    sums = np.zeros(3, dtype=np.int64)
    w_col = np.arange(1, COLS + 1, dtype=np.int64)[None, :]
    for s0 in range(i0, min(ROWS, i0 + BLOCK), SUB):
        s1 = min(ROWS, i0 + BLOCK, s0 + SUB)
        rows = np.arange(s0, s1, dtype=np.uint64)
        idx = cols[None, :] * np.uint64(ROWS) + rows[:, None]
        gd = rng.gaussian_at(SEED, 0, idx) * d[None, :]
        bf = torch.from_numpy(gd).to(torch.bfloat16)
        _shared[s0:s1] = bf.to(torch.float32).numpy()
        bits = bf.view(torch.int16).numpy().astype(np.int64) & 0xFFFF
        w_row = (np.arange(s0, s1, dtype=np.int64) % 1009 + 1)[:, None]
        sums += np.array([bits.sum(), (bits * w_col).sum(), (bits * w_row).sum()], dtype=np.int64)
    return i0, sums


def main() -> None:
    global _shared
    from oracle import rng
    t_start = time.time()
    buf = mmap.mmap(-1, ROWS * COLS * 4)
    _shared = np.frombuffer(buf, dtype=np.float32).reshape(ROWS, COLS)
    xs = rng.gaussian_at(SEED, 1, np.arange(COLS, dtype=np.uint64))
    e = rng.gaussian_at(SEED, 2, np.arange(ROWS, dtype=np.uint64))
    starts = list(range(0, ROWS, BLOCK))
    checks = np.zeros(3, dtype=np.int64)
    R = None
    ax = np.empty(ROWS)
    ctx = mp.get_context("fork")
    with ctx.Pool(6) as pool:
        for i0, sums in pool.imap(_block, starts):
            checks += sums
            i1 = min(ROWS, i0 + BLOCK)
            a = _shared[i0:i1].astype(np.float64)
            ax[i0:i1] = a @ xs
            aug = np.empty((i1 - i0, COLS + 2))
            aug[:, :COLS] = a
            aug[:, COLS] = ax[i0:i1]
            aug[:, COLS + 1] = e[i0:i1]
            stacked = aug if R is None else np.vstack([R, aug])
            R = np.linalg.qr(stacked, mode="r")
            print(f"rows {i1}: {time.time() - t_start:.0f} s", flush=True)
    n = COLS
    nax, ne = np.linalg.norm(ax), np.linalg.norm(e)
    coef = NOISE * nax / ne
    b = ax + coef * e
    qtb = R[:n, n] + coef * R[:n, n + 1]
    R11 = R[:n, :n]
    import scipy.linalg as sl
    x = sl.solve_triangular(R11, qtb)
    # LS residual norm from the augmented factor: the QR of [A | b] has
    # |R_{n+1,n+1}| = ||r_LS||; here b = A x* + coef e, so form the 2x2 tail.
    tail = R[n:n + 2, n:n + 2]
    r_ls_qr = float(np.linalg.norm(tail @ np.array([1.0, coef])))
    # second pass over the stored A: explicit residual, gradient, CSNE step
    r = np.empty(ROWS)
    g = np.zeros(n)
    fro2 = 0.0
    for i0 in range(0, ROWS, BLOCK):
        i1 = min(ROWS, i0 + BLOCK)
        a = _shared[i0:i1].astype(np.float64)
        r[i0:i1] = b[i0:i1] - a @ x
        g += a.T @ r[i0:i1]
        fro2 += float(np.einsum("ij,ij->", a, a))
    nr = float(np.linalg.norm(r))
    normal_residual = float(np.linalg.norm(g) / (np.sqrt(fro2) * nr))
    delta = sl.solve_triangular(R11, sl.solve_triangular(R11, g, trans="T"))
    csne_rel = float(np.linalg.norm(delta) / np.linalg.norm(x))
    sv = np.linalg.svd(R11, compute_uv=False)
    out = dict(
        x_ls=x, x_csne=x + delta, x_star=xs,
        bits_checksums=checks,
        rows=np.int64(ROWS), cols=np.int64(COLS), decades=np.float64(DECADES), seed=np.int64(SEED),
        noise=np.float64(NOISE),
        b_norm=np.float64(np.linalg.norm(b)), ax_norm=np.float64(nax), residual_norm=np.float64(nr),
        residual_norm_qr=np.float64(r_ls_qr), normal_residual=np.float64(normal_residual),
        csne_relative_step=np.float64(csne_rel), frobenius=np.float64(np.sqrt(fro2)),
        sigma_max=np.float64(sv[0]), sigma_min=np.float64(sv[-1]),
    )
    np.savez_compressed(OUT, **out)
    print(OUT)
    for k in ("b_norm", "ax_norm", "residual_norm", "residual_norm_qr", "normal_residual", "csne_relative_step",
              "frobenius", "sigma_max", "sigma_min"):
        print(f"  {k} = {float(out[k]):.6e}")
    print(f"  ||x_ls - x*|| / ||x*|| = {np.linalg.norm(x - xs) / np.linalg.norm(xs):.3e}")
    print(f"  checksums = {checks.tolist()}  ({time.time() - t_start:.0f} s)")


if __name__ == "__main__":
    main()
