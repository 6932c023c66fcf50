"""Pins the oracle's operators, workspace, validator and drivers against
closed forms, numpy/scipy LAPACK and textbook LSQR."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import operators as O
from oracle import results as R
from oracle import solvers as S
from oracle import validation as V
from oracle.clock import SyntheticClock, WallClock
from oracle.workspace import Workspace


def small_dense(m=120, n=30, decades=2.0, seed=0, noise=1e-3):
    r = np.random.default_rng(seed)
    a = r.standard_normal((m, n)) * np.logspace(0, -decades, n)
    b = a @ r.standard_normal(n) + noise * r.standard_normal(m)
    return a, b


def test_readme_diagonal_example():
    # README.md:534-583: A = diag(100, 3, 2, 1), b = (100, 6, 6, 4) -> x = (1, 2, 3, 4)
    op = O.DiagonalOperator([100.0, 3.0, 2.0, 1.0])
    r = S.solve_amfls(op, np.array([100.0, 6.0, 6.0, 4.0]), tolerance=1e-10, seed=7)
    assert r.status == R.SUCCESS
    assert np.allclose(r.solution, [1, 2, 3, 4], rtol=0, atol=1e-12)


def test_operator_products_and_norm_bounds():
    r = np.random.default_rng(3)
    a = r.standard_normal((40, 15))
    a[np.abs(a) < 0.8] = 0.0
    import scipy.sparse as sp
    s = sp.csr_matrix(a)
    csr = O.CsrOperator(40, 15, s.indptr, s.indices, s.data)
    for op in (O.DenseOperator(a), O.RowMajorDenseOperator(a), csr):
        x = r.standard_normal((15, 3))
        assert np.allclose(op.apply(x), a @ x, rtol=1e-14, atol=1e-14)
        assert np.allclose(op.apply_transpose(a @ x), a.T @ (a @ x), rtol=1e-14, atol=1e-13)
        up, lo, sa, st = op.validation_error_model()
        true = np.linalg.norm(a, 2)
        assert lo <= true <= up
        assert up >= np.linalg.norm(np.abs(a), 2)
    up, lo, sa, st = csr.validation_error_model()
    nnz_rows = np.count_nonzero(a, axis=1).max()
    nnz_cols = np.count_nonzero(a, axis=0).max()
    assert (sa, st) == (2 * nnz_rows + 2, 2 * nnz_cols + 2)
    # rank-one closed form: ||u v^T|| = ||u|| ||v|| and the Frobenius bound is tight
    u, v = r.standard_normal(7), r.standard_normal(5)
    up, lo, _, _ = O.DenseOperator(np.outer(u, v)).validation_error_model()
    assert lo <= np.linalg.norm(u) * np.linalg.norm(v) <= up <= np.linalg.norm(u) * np.linalg.norm(v) * (1 + 1e-12)


def test_csr_constructor_rejects_bad_structure():
    with pytest.raises(ValueError):
        O.CsrOperator(2, 2, [0, 2, 2], [0, 0], [1.0, 2.0])  # duplicate column in a row
    with pytest.raises(ValueError):
        O.CsrOperator(2, 2, [0, 1, 2], [0, 5], [1.0, 2.0])  # out of range
    with pytest.raises(ValueError):
        O.CsrOperator(2, 2, [0, 1, 2], [0, 1], [1.0, np.inf])


def test_long_double_refinement_is_exact_on_small_integers():
    a = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])
    op = O.DenseOperator(a)
    y = op.apply_validation_refinement(np.array([1.0, 1.0], dtype=np.longdouble))
    assert list(y) == [3.0, 7.0, 11.0]


def test_validation_bounds_enclose_true_backward_error():
    a, b = small_dense()
    op = O.DenseOperator(a)
    stats = R.RunStatistics()
    c = V.Counted(op, stats, WallClock())
    xo = np.linalg.lstsq(a, b, rcond=None)[0]
    x = xo + 1e-6 * np.random.default_rng(5).standard_normal(xo.size)
    res = V.validate_original_candidate(c, b, x, 0.0, 1e-8)
    r = a @ x - b
    nrm = np.linalg.norm(a, 2)
    true_ls = np.linalg.norm(a.T @ r) / (nrm * np.linalg.norm(r))
    true_comp = np.linalg.norm(r) / (nrm * np.linalg.norm(x) + np.linalg.norm(b))
    assert res.least_squares_backward_error_upper_bound >= true_ls
    assert res.compatible_backward_error_upper_bound >= true_comp
    # an exact solution: the nominal contract sees exact stationarity, the
    # certified one keeps the callback radius and certifies the compatible
    # backward error instead (candidate_validation.cpp:694-814)
    opd = O.DiagonalOperator([2.0, 4.0])
    cd = V.Counted(opd, R.RunStatistics(), WallClock())
    rd = V.validate_original_candidate(cd, np.array([2.0, 4.0]), np.array([1.0, 1.0]), 0.0, 1e-8)
    assert rd.normal_residual_norm == 0.0
    assert rd.status == R.SUCCESS and rd.stop_reason == R.COMPATIBLE_BACKWARD_ERROR
    assert rd.compatible_backward_error_upper_bound < 1e-14


def test_ridge_energy_bound_encloses_truth():
    a, b = small_dense(seed=4)
    lam = 1e-2
    op = O.DenseOperator(a)
    xr = np.linalg.solve(a.T @ a + lam * np.eye(a.shape[1]), a.T @ b)
    x = xr + 1e-5 * np.random.default_rng(9).standard_normal(xr.size)
    c = V.Counted(op, R.RunStatistics(), WallClock())
    res = V.validate_original_candidate(c, b, x, lam, 1e-8)
    e = x - xr
    true_energy = math.sqrt(np.linalg.norm(a @ e) ** 2 + lam * np.linalg.norm(e) ** 2)
    assert res.energy_error_upper_bound >= true_energy


def test_workspace_krylov_invariants():
    a, b = small_dense(m=80, n=25, seed=6)
    op = O.DenseOperator(a)
    ws = Workspace(op, b, 0.0, 0, R.RunStatistics(), SyntheticClock())
    ws.initialize(np.random.default_rng(1).standard_normal((80, 2)))
    for _ in range(5):
        assert ws.advance_level()
    Vb = ws.V
    assert np.allclose(Vb.T @ Vb, np.eye(Vb.shape[1]), atol=1e-12)
    # nested objective decrease of the projected candidates
    prev = math.inf
    T = ws.projected_operator()
    g = np.array(ws.g)
    for k in range(1, T.shape[1] + 1):
        y = np.linalg.lstsq(T[:, :k], g, rcond=None)[0]
        obj = np.linalg.norm(T[:, :k] @ y - g)
        assert obj <= prev + 1e-12
        prev = obj
    # the candidate is the projected LS solution: it equals the LS solution
    # over span(V) computed directly from A
    x = ws.projected_solution()
    yv = np.linalg.lstsq(a @ Vb, b, rcond=None)[0]
    assert np.allclose(x, Vb @ yv, rtol=1e-9, atol=1e-9)


def test_amfls_matches_lstsq_and_ridge_closed_form():
    a, b = small_dense(m=300, n=60, decades=3.0, seed=8, noise=1e-4)
    op = O.DenseOperator(a)
    r = S.solve_amfls(op, b, tolerance=1e-10, clock=SyntheticClock())
    xo = np.linalg.lstsq(a, b, rcond=None)[0]
    assert np.linalg.norm(r.solution - xo) / np.linalg.norm(xo) < 1e-6
    lam = 1e-3
    xr = np.linalg.solve(a.T @ a + lam * np.eye(60), a.T @ b)
    r = S.solve_amfls(op, b, regularization=lam, tolerance=1e-10, clock=SyntheticClock())
    assert r.status == R.SUCCESS and r.stop_reason == R.RELATIVE_ENERGY_ERROR
    assert np.linalg.norm(r.solution - xr) / np.linalg.norm(xr) < 1e-8
    # counter identity: one A* call per level plus validations
    s = r.statistics
    assert s.search_at_block_calls == s.search_a_block_calls
    assert s.a_columns == s.search_a_columns + s.validation_a_columns


def test_lsqr_matches_scipy_textbook_lsqr():
    a, b = small_dense(m=200, n=40, decades=1.0, seed=10)
    op = O.DenseOperator(a)
    for lam in (0.0, 1e-2):
        r = S.solve_lsqr(op, b, regularization=lam, tolerance=1e-12, maximum_iterations=25)
        ref = spla.lsqr(a, b, damp=math.sqrt(lam), atol=0, btol=0, conlim=0, iter_lim=r.depth)[0]
        assert np.linalg.norm(r.solution - ref) / np.linalg.norm(ref) < 1e-10


def test_fixed_rbgk_deterministic_and_statuses():
    a, b = small_dense(m=90, n=30, seed=12)
    op = O.DenseOperator(a)
    r1 = S.solve_fixed_rbgk(op, b, auxiliary_width=2, maximum_depth=4, tolerance=1e-14, seed=3)
    r2 = S.solve_fixed_rbgk(op, b, auxiliary_width=2, maximum_depth=4, tolerance=1e-14, seed=3)
    assert np.array_equal(r1.solution, r2.solution)
    assert r1.status == R.WORK_LIMIT and r1.stop_reason == R.MAXIMUM_DEPTH
    r3 = S.solve_amfls(op, b, tolerance=1e-14, maximum_basis_size=7, clock=SyntheticClock())
    assert r3.status == R.BASIS_LIMIT and r3.basis_rank == 7


def test_nan_operator_reports_numerical_breakdown():
    class NanOp(O.DenseOperator):
        def apply(self, x):
            y = super().apply(x)
            y[0] = np.nan
            return y
    a, b = small_dense(m=20, n=5, seed=2)
    r = S.solve_amfls(NanOp(a), b, tolerance=1e-8, clock=SyntheticClock())
    assert r.status == R.NUMERICAL_BREAKDOWN


def test_result_tensor_layout_widths():
    a, b = small_dense(m=30, n=6, seed=3)
    r = S.solve_amfls(O.DenseOperator(a), b, tolerance=1e-10, clock=SyntheticClock())
    sol, met, cnt, tim, tf, ti, st, sr = R.to_tensors(r)
    assert met.shape == (21,) and cnt.shape == (36,) and tim.shape == (9,)
    assert tf.shape == (len(r.trace), 29) and ti.shape == (len(r.trace), 4)
    assert met[10] == r.backward_error_upper_bound and cnt[11] == r.basis_rank


def test_rank_deficient_operator_stops_at_rank():
    # A = P Q^T of rank 3: the Krylov basis cannot exceed 3 columns and the
    # certified answer is the minimum-norm least-squares solution.
    r = np.random.default_rng(21)
    a = r.standard_normal((60, 3)) @ r.standard_normal((3, 12))
    b = r.standard_normal(60)
    res = S.solve_amfls(O.DenseOperator(a), b, tolerance=1e-10, clock=SyntheticClock())
    assert res.basis_rank <= 3
    xo = np.linalg.lstsq(a, b, rcond=None)[0]
    assert np.linalg.norm(res.solution - xo) / np.linalg.norm(xo) < 1e-8
