"""Writes tests/golden/fixtures.npz: inputs and oracle outputs for the GPU
parity tests (run from the repository root: python tests/golden/make_fixtures.py).

The expected outputs come from the pinned CPU restatement in oracle/ (the
reference's own C++ needs an external BLAS/LAPACK and is not built here), with
the synthetic clock so the adaptive width schedule is deterministic.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import operators as O  # noqa: E402
from oracle import rng  # noqa: E402
from oracle import solvers as S  # noqa: E402
from oracle.clock import SyntheticClock  # noqa: E402
from amfls_solver_amd import problems as P  # noqa: E402


def trace_array(res):
    return np.array([[t.depth, t.basis_rank, t.auxiliary_width, t.backward_error_upper_bound,
                      t.relative_energy_error_upper_bound, t.contract_status] for t in res.trace])


def main():
    out = {}
    out["gauss_block"] = rng.gaussian_matrix(64, 4, 7, 3, 2)
    a, b, _ = P.dense_gd(300, 80, 3.0, seed=2, noise=1e-4)
    out["dense_a"], out["dense_b"] = a, b
    op = O.DenseOperator(a)
    r = S.solve_fixed_rbgk(op, b, auxiliary_width=3, maximum_depth=40, tolerance=1e-10, seed=5,
                           clock=SyntheticClock())
    out["rbgk_solution"], out["rbgk_trace"] = r.solution, trace_array(r)
    out["rbgk_status"] = np.array([r.status, r.stop_reason, r.depth, r.basis_rank])
    for tag, lam in (("l0", 0.0), ("l3", 1e-3)):
        r = S.solve_amfls(op, b, regularization=lam, tolerance=1e-10, seed=11, clock=SyntheticClock())
        out[f"amfls_dense_{tag}_solution"] = r.solution
        out[f"amfls_dense_{tag}_trace"] = trace_array(r)
        out[f"amfls_dense_{tag}_status"] = np.array([r.status, r.stop_reason, r.depth, r.basis_rank,
                                                     r.auxiliary_width])
    p = P.csr_rows(3000, 700, 9, seed=5)
    out["csr_rp"], out["csr_ci"], out["csr_val"], out["csr_b"] = p.row_ptr, p.col_idx, p.values, p.b
    cop = O.CsrOperator(p.rows, p.cols, p.row_ptr, p.col_idx, p.values.astype(np.float64))
    out["csr_model"] = np.array(cop.validation_error_model(), dtype=np.float64)
    for tag, lam in (("l0", 0.0), ("l3", 1e-3)):
        r = S.solve_amfls(cop, p.b, regularization=lam, tolerance=1e-9, seed=1, clock=SyntheticClock())
        out[f"amfls_csr_{tag}_solution"] = r.solution
        out[f"amfls_csr_{tag}_trace"] = trace_array(r)
        out[f"amfls_csr_{tag}_status"] = np.array([r.status, r.stop_reason, r.depth, r.basis_rank,
                                                   r.auxiliary_width])
    r = S.solve_lsqr(cop, p.b, regularization=1e-3, tolerance=1e-9)
    out["lsqr_csr_solution"], out["lsqr_csr_status"] = r.solution, np.array([r.status, r.depth])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
