"""The engine's certified validator against the oracle's on the same
candidate (amfls_validate_candidate vs oracle validate_original_candidate,
reference candidate_validation.cpp:1176-1378).

Away from convergence no bound is dominated by cancellation, so the two
implementations (device FP64 products vs numpy) must give every bound to
1e-12 relative.  The forced long-double refinement and the forced one-step
ridge correction must take the same branch with the same disposition and
operator accounting, their bounds agreeing to 1e-9.  On a converged
candidate, where the normal residual is near rounding level, the bounds
agree to 1e-6 and both enclose the true LS backward error."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")
BOUNDS = ["compatible_backward_error_upper_bound", "least_squares_backward_error_upper_bound",
          "backward_error_upper_bound", "relative_normal_residual_upper_bound", "energy_error_upper_bound",
          "relative_energy_error_upper_bound", "objective_gap_upper_bound", "operator_norm_lower_bound",
          "augmented_operator_norm_lower_bound"]
NORMS = ["residual_norm", "solution_norm", "augmented_residual_norm"]
RIDGE = ["ridge_base_energy_error_upper_bound", "ridge_corrected_energy_error_upper_bound"]


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


@pytest.fixture(scope="module")
def torch_mod(gpu):
    import torch
    return torch


def _csr(torch, rp, ci, val, shape):
    return torch.sparse_csr_tensor(torch.from_numpy(np.asarray(rp, np.int64)),
                                   torch.from_numpy(np.asarray(ci, np.int64)),
                                   torch.from_numpy(np.asarray(val)), size=shape)


def _oracle(op, b, x, lam, tol):
    from oracle import clock as CL, results as R, validation as V
    c = V.Counted(op, R.RunStatistics(), CL.SyntheticClock())
    return V.validate_original_candidate(c, b, x, lam, tol)


def _same(name, e, o, rtol):
    if e is None or o is None:
        assert e is None and o is None or (e is None and np.isnan(o)) or (o is None and np.isnan(e)), name
        return
    if not np.isfinite(o) or not np.isfinite(e):
        assert (np.isnan(o) and np.isnan(e)) or e == o, (name, e, o)
        return
    assert abs(e - o) <= rtol * max(abs(o), 1e-300), (name, e, o, abs(e - o) / max(abs(o), 1e-300))


def _compare(eng, ora, rtol, fields):
    assert (int(eng.status), int(eng.stop_reason)) == (int(ora.status), int(ora.stop_reason))
    for f in fields:
        _same(f, getattr(eng, f), getattr(ora, f), rtol)


def _problems(torch, fx):
    from oracle import operators as O
    A = fx["dense_a"]
    yield "dense", torch.from_numpy(A), O.DenseOperator(A), fx["dense_b"], fx["amfls_dense_l0_solution"], None
    val = fx["csr_val"]
    cop = O.CsrOperator(3000, 700, fx["csr_rp"], fx["csr_ci"], val.astype(np.float64))
    yield ("csr", _csr(torch, fx["csr_rp"], fx["csr_ci"], val, (3000, 700)), cop, fx["csr_b"],
           fx["amfls_csr_l0_solution"], torch.float32)


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_unconverged_candidate_bounds_match_oracle(torch_mod, fx, lam):
    import amfls_solver_amd as am
    torch = torch_mod
    rng = np.random.default_rng(21)
    for kind, mat, op, b, sol, vdt in _problems(torch, fx):
        x = 0.5 * sol + 0.01 * rng.standard_normal(sol.size) * np.linalg.norm(sol) / np.sqrt(sol.size)
        eng = am.validate_candidate(mat, torch.from_numpy(b), torch.from_numpy(x), regularization=lam,
                                    tolerance=1e-8, value_dtype=vdt)
        ora = _oracle(op, b, x, lam, 1e-8)
        _compare(eng, ora, 1e-12, BOUNDS + NORMS + ["normal_residual_norm"] + (RIDGE if lam > 0 else []))
        assert eng.ridge_correction_disposition == ora.ridge_correction_disposition, kind
        assert np.array_equal(eng.solution.numpy(), x)


def test_converged_candidate_bounds_match_and_enclose(torch_mod, fx):
    import amfls_solver_amd as am
    torch = torch_mod
    for kind, mat, op, b, sol, vdt in _problems(torch, fx):
        eng = am.validate_candidate(mat, torch.from_numpy(b), torch.from_numpy(sol), tolerance=1e-10,
                                    value_dtype=vdt)
        ora = _oracle(op, b, sol, 0.0, 1e-10)
        _compare(eng, ora, 1e-6, ["compatible_backward_error_upper_bound", "operator_norm_lower_bound"] + NORMS)
        # the normal residual sits near rounding level: both bounds enclose the truth
        if kind == "dense":
            A64 = fx["dense_a"]
        else:
            import scipy.sparse as sp
            A64 = sp.csr_matrix((fx["csr_val"].astype(np.float64), fx["csr_ci"], fx["csr_rp"]),
                                shape=(3000, 700)).toarray()
        r = A64 @ sol - b
        g = A64.T @ r
        smax = np.linalg.svd(A64, compute_uv=False)[0]
        true_ls = np.linalg.norm(g) / (smax * (1 + 1e-12) * np.linalg.norm(r))
        assert eng.least_squares_backward_error_upper_bound >= true_ls, kind
        assert ora.least_squares_backward_error_upper_bound >= true_ls, kind


def test_forced_long_double_refinement_matches_oracle(torch_mod):
    """Rounded exact LS solution of a 6 x 3 integer problem at an
    unreachable tolerance: the FP64 gradient is below its error radius, so
    both validators run the long-double branch (two extra products)."""
    import amfls_solver_amd as am
    from oracle import operators as O
    torch = torch_mod
    A = np.array([[1, 2, 0], [0, 1, 1], [3, 0, 1], [1, 1, 1], [2, -1, 0], [0, 0, 4]], dtype=np.float64)
    b = np.array([1, -2, 3, 0, 5, 1], dtype=np.float64)
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    eng = am.validate_candidate(torch.from_numpy(A), torch.from_numpy(b), torch.from_numpy(x), tolerance=1e-30)
    ora = _oracle(O.DenseOperator(A), b, x, 0.0, 1e-30)
    assert eng.base_validation_a_columns == ora.statistics.base_validation_a_columns == 2
    assert eng.base_validation_at_columns == ora.statistics.base_validation_at_columns == 2
    _compare(eng, ora, 1e-9, ["compatible_backward_error_upper_bound", "least_squares_backward_error_upper_bound",
                              "backward_error_upper_bound", "operator_norm_lower_bound"])


@pytest.mark.parametrize("layout", ["column_major_f64", "padded_row_major_f32"])
def test_forced_refinement_streams_large_dense_matrix(torch_mod, layout):
    """The long-double refinement streams A to the host in chunks of 2^22
    entries: a 1.5e6 x 3 integer problem spans two chunks (1398101 rows and
    the rest).  Column-major storage (a transposed view) and row-major
    storage with a padded leading dimension take the two copy shapes; both
    validators must take the refinement branch and agree with the oracle."""
    import amfls_solver_amd as am
    from oracle import operators as O
    torch = torch_mod
    rng = np.random.default_rng(7)
    m, n = 1_500_000, 3
    A = rng.integers(-4, 5, size=(m, n)).astype(np.float64)
    b = rng.integers(-4, 5, size=m).astype(np.float64)
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    if layout == "column_major_f64":
        At = torch.from_numpy(np.ascontiguousarray(A.T)).to("cuda").t()
        assert At.stride() == (1, m)
    else:
        At = torch.zeros(m, n + 5, dtype=torch.float32, device="cuda")
        At[:, :n] = torch.from_numpy(A.astype(np.float32)).to("cuda")
        At = At[:, :n]
        assert At.stride() == (n + 5, 1)
    eng = am.validate_candidate(At, torch.from_numpy(b), torch.from_numpy(x), tolerance=1e-30)
    ora = _oracle(O.DenseOperator(A), b, x, 0.0, 1e-30)
    assert eng.base_validation_a_columns == ora.statistics.base_validation_a_columns == 2
    assert eng.base_validation_at_columns == ora.statistics.base_validation_at_columns == 2
    _compare(eng, ora, 1e-9, ["compatible_backward_error_upper_bound", "least_squares_backward_error_upper_bound",
                              "backward_error_upper_bound", "operator_norm_lower_bound"])


def test_forced_ridge_correction_matches_oracle(torch_mod):
    """A = diag(10, 1), lambda = 1e-2, an error along the first eigenvector:
    both validators attempt the one-step ridge correction, improve the
    bound and agree on it."""
    import amfls_solver_amd as am
    from oracle import operators as O, results as R
    torch = torch_mod
    lam = 1e-2
    A = np.diag([10.0, 1.0])
    b = np.array([1.0, 1.0])
    x = np.array([10.0 / 100.01 + 1e-6, 1.0 / 1.01])
    eng = am.validate_candidate(torch.from_numpy(A), torch.from_numpy(b), torch.from_numpy(x), regularization=lam,
                                tolerance=1e-12)
    ora = _oracle(O.DenseOperator(A), b, x, lam, 1e-12)
    assert eng.ridge_correction_disposition == ora.ridge_correction_disposition == R.RC_ATTEMPTED_IMPROVED
    assert eng.ridge_correction_a_columns == ora.statistics.ridge_correction_a_columns == 1
    _compare(eng, ora, 1e-9, RIDGE + ["energy_error_upper_bound", "relative_energy_error_upper_bound",
                                      "ridge_correction_gamma"])
