"""Hand-computed cases that pin the oracle's controller state machine and
its validator's rounding radii, operator models, long-double refinement
branch and one-step ridge correction (CPU only).

Every expected value below is written out from the rules the oracle
restates (reference amfls_internal.hpp:35-674 and
candidate_validation.cpp:96-119, 395-566, 822-926, 951-1149, 1176-1378),
worked by hand in the comments, or computed exactly with rationals; none
is produced by calling the function under test a second time."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import clock as CL
from oracle import controller as C
from oracle import operators as O
from oracle import results as R
from oracle import validation as V

U = Fraction(1, 2 ** 53)  # binary64 unit roundoff
LN10 = math.log(10.0)


def _history(op_level, loc_level):
    """Width-1 history: 2 levels from 1 to 1e-2 (ln 100 = 4.60517), then 2
    levels from 1e-2 to 1e-2 * 10^-0.9 (0.9 ln 10 = 2.07233)."""
    h = C.ProgressHistory()
    c2 = 1e-2 * 10 ** -0.9
    C.record_deepen_progress(h, 1.0, 1e-2, 1, 1, 2, 2 * op_level, 2 * loc_level)
    C.record_deepen_progress(h, 1e-2, c2, 1, 1, 2, 2 * op_level, 2 * loc_level)
    return h, c2


def test_progress_history_records_two_intervals_then_pools():
    h, c2 = _history(1e-3, 1e-4)
    assert h.active_left_width == 1 and h.same_width_level_interval_count == 2
    assert h.same_width_interval_levels == [2, 2]
    assert h.same_width_log_contractions == pytest.approx([2 * LN10, 0.9 * LN10], rel=1e-14)
    assert h.same_width_level_log_contractions == pytest.approx([LN10, 0.45 * LN10], rel=1e-14)
    assert h.same_width_level_operator_seconds == pytest.approx([1e-3, 1e-3], rel=1e-15)
    assert h.same_width_level_local_seconds == pytest.approx([1e-4, 1e-4], rel=1e-15)
    assert h.prior_same_width_levels == 0
    # a third interval shifts the window and pools the oldest into the prior
    C.record_deepen_progress(h, c2, c2 / 10.0, 1, 1, 1, 1e-3, 1e-4)
    assert h.prior_same_width_levels == 2
    assert h.prior_same_width_log_contraction == pytest.approx(2 * LN10, rel=1e-14)
    assert h.same_width_interval_levels == [2, 1]
    assert h.same_width_log_contractions == pytest.approx([0.9 * LN10, LN10], rel=1e-13)
    # a width change resets the history to the new width
    C.record_deepen_progress(h, 1.0, 0.5, 2, 2, 1, 1e-3, 1e-4)
    assert h.active_left_width == 2 and h.same_width_level_interval_count == 1 and h.prior_same_width_levels == 0


def test_initial_width_candidate_scalar_slowdown_widens():
    # recent rate (4.60517 + 2.07233) / 4 = 1.66937 per level; remaining
    # ln(c2 / 1e-8) = 5.1 ln 10 = 11.7432 -> 7.03 levels, forecast 8;
    # target width 2 can run min(100 - 10, (1000 - 10) // 2) = 90 levels;
    # horizon 1 * 8 = 8 columns: 8 levels at width 1, 4 at width 2;
    # seconds: incumbent 8 * (1e-3 + 1e-4) = 8.8e-3, target (local x 2^2)
    # 4 * (1e-3 + 4e-4) + 4e-4 = 6.0e-3 -> the target is cheaper;
    # level rates 2.3026 then 1.0362 < 2.3026 / 2: a recent slowdown, which
    # is what a scalar frontier needs to widen.
    h, c2 = _history(1e-3, 1e-4)
    d = C.make_initial_width_candidate(1, c2, 1e-8, 10, 100, 10, 1000, 1, h)
    assert d.finite_current_certificate and d.history_mature and not d.progress_stalled
    assert d.recent_level_log_contraction == pytest.approx(1.6693741924206833, rel=1e-13)
    assert d.remaining_log_contraction == pytest.approx(5.1 * LN10, rel=1e-13)
    assert d.forecast_levels == 8
    assert d.target_remaining_levels == 90 and d.matched_horizon == 8
    assert (d.incumbent_levels, d.target_levels) == (8, 4)
    assert d.incumbent_estimated_seconds == pytest.approx(8.8e-3, rel=1e-12)
    assert d.target_estimated_seconds == pytest.approx(6.0e-3, rel=1e-12)
    assert d.pass_reduction_admissible and d.recent_slowdown and not d.long_horizon_pressure
    assert not d.prior_history_mature and not d.persistent_slowdown
    assert d.widening_necessary and d.candidate


def test_initial_width_candidate_rejects_costlier_target():
    # same progress, but per-level local work 1e-3 dominates operator 1e-4:
    # incumbent 8 * 1.1e-3 = 8.8e-3, target 4 * (1e-4 + 4e-3) + 4e-3 = 2.04e-2
    h, c2 = _history(1e-4, 1e-3)
    d = C.make_initial_width_candidate(1, c2, 1e-8, 10, 100, 10, 1000, 1, h)
    assert d.target_estimated_seconds == pytest.approx(2.04e-2, rel=1e-12)
    assert not d.pass_reduction_admissible and d.widening_necessary and not d.candidate


def test_initial_width_candidate_block_stall_widens():
    # width 2, two intervals of 3 levels without progress: stalled;
    # target 4 fits min(90, (200 - 100) // 4) = 25 levels -> horizon 100;
    # 50 levels at width 2 against 25 at width 4: incumbent 50 * 1.1e-3 =
    # 0.055 s, target 25 * (1e-3 + 4e-4) + 4e-4 = 0.0354 s; a stalled block
    # frontier under long-horizon pressure widens.
    h = C.ProgressHistory()
    for _ in range(2):
        C.record_deepen_progress(h, 1e-3, 1e-3, 2, 2, 3, 3e-3, 3e-4)
    d = C.make_initial_width_candidate(2, 1e-3, 1e-8, 10, 100, 100, 200, 2, h)
    assert d.progress_stalled and d.recent_level_log_contraction == 0.0 and d.forecast_levels == 0
    assert (d.matched_horizon, d.incumbent_levels, d.target_levels) == (100, 50, 25)
    assert d.incumbent_estimated_seconds == pytest.approx(0.055, rel=1e-12)
    assert d.target_estimated_seconds == pytest.approx(0.0354, rel=1e-12)
    assert d.long_horizon_pressure and not d.recent_slowdown and d.widening_necessary and d.candidate


def test_initial_width_candidate_guards():
    h, c2 = _history(1e-3, 1e-4)
    # certificate already within tolerance
    d = C.make_initial_width_candidate(1, 1e-9, 1e-8, 10, 100, 10, 1000, 1, h)
    assert not d.finite_current_certificate and not d.candidate and d.forecast_levels == 0
    # history recorded at another width
    d = C.make_initial_width_candidate(1, c2, 1e-8, 10, 100, 10, 1000, 2, h)
    assert not d.history_mature and not d.candidate
    # one interval only
    h1 = C.ProgressHistory()
    C.record_deepen_progress(h1, 1.0, 1e-2, 1, 1, 2, 2e-3, 2e-4)
    assert not C.make_initial_width_candidate(1, c2, 1e-8, 10, 100, 10, 1000, 1, h1).candidate


def test_matched_horizon_plan_lifecycle():
    p = C.MatchedHorizonPlan()
    assert C.remaining_matched_horizon(p, 5) == 0
    assert C.start_matched_horizon_plan(p, 10, 100, 30, 2, 4, 1e-3)
    assert (p.end_basis_rank, p.origin_active_width, p.active_left_width) == (40, 2, 4)
    assert C.remaining_matched_horizon(p, 25) == 15 and C.remaining_matched_horizon(p, 40) == 0
    assert not C.extend_matched_horizon_to_basis_limit(p, 35, 100)        # horizon not yet reached
    assert C.extend_matched_horizon_to_basis_limit(p, 40, 100) and p.end_basis_rank == 100
    # stalled: less than a halving since the start (ln(1e-3/6e-4) = 0.51 < ln 2)
    assert C.matched_horizon_progress_stalled(p, 6e-4)
    assert not C.matched_horizon_progress_stalled(p, 4e-4)               # ln 2.5 = 0.92
    C.reset_matched_horizon_costs(p, 4)
    assert C.record_matched_horizon_cost(p, 4, 4, 2, 0.02, 0.01)
    assert p.level_operator_seconds[0] == pytest.approx(0.01) and p.level_local_seconds[0] == pytest.approx(0.005)
    assert not C.record_matched_horizon_cost(p, 4, 8, 2, 0.02, 0.01)       # width changed inside
    assert not C.record_matched_horizon_cost(p, 2, 2, 2, 0.02, 0.01)       # not the plan's width
    assert C.record_matched_horizon_cost(p, 4, 4, 1, 0.003, 0.001)
    assert not C.record_matched_horizon_cost(p, 4, 4, 1, 0.003, 0.001)     # two samples at most
    # a horizon beyond the basis limit resets the plan
    assert not C.start_matched_horizon_plan(p, 10, 20, 30, 2, 4, 1e-3)
    assert p.end_basis_rank == -1 and C.remaining_matched_horizon(p, 10) == 0


def _within_steps_above(value: float, exact: Fraction, steps: int) -> bool:
    v = Fraction(value)
    if v < exact:
        return False
    hi = value
    for _ in range(steps):
        hi = float(np.nextafter(hi, math.inf))
    return exact <= v <= Fraction(hi)


def test_callback_error_radius_hand_values():
    denorm = Fraction(2) ** -1074
    model = (2.0, 1.0, 4, 6)
    # A x with ||x|| <= 3 on an operator whose products take 4 roundings:
    # gamma_4 * 2 * 3 + 4 sqrt(5) denorm_min (5 outputs)
    r = V.callback_error_radius(model, False, 3.0, 5)
    main = 6 * (4 * U / (1 - 4 * U))
    assert Fraction(r) >= main + 4 * Fraction(2236, 1000) * denorm
    assert Fraction(r) <= main * (1 + 16 * U) + 5 * Fraction(224, 100) * denorm
    # A^T y uses the transpose step count
    rt = V.callback_error_radius(model, True, 3.0, 5)
    # (gamma and the product are each stepped upward: a few units of roundoff above the exact value)
    assert 6 * (6 * U / (1 - 6 * U)) <= Fraction(rt) <= 6 * (6 * U / (1 - 6 * U)) * (1 + 16 * U) + 20 * denorm
    assert V.callback_error_radius(model, False, 0.0, 5) == 0.0
    assert V.callback_error_radius((math.inf, 0.0, 4, 6), False, 3.0, 5) == math.inf
    assert V.callback_error_radius((2.0, 3.0, 4, 6), False, 3.0, 5) == math.inf       # lower > upper
    # one rounded operation on a stored value of norm 10: gamma_2 * 10
    one = V.one_operation_radius(10.0, False, 3)
    assert _within_steps_above(one, 10 * (2 * U / (1 - 2 * U)), 3)


def test_operator_models_hand_values():
    # [[3, 0], [4, 0]] is rank one with sigma = 5: Frobenius 5, sqrt(7 * 4),
    # largest column 2-norm 5 -> both bounds are 5; dense steps 2n+2, 2m+2
    up, lo, sa, st = O.DenseOperator(np.array([[3.0, 0.0], [4.0, 0.0]])).validation_error_model()
    assert (sa, st) == (6, 6)
    assert 5.0 <= up <= 5.0 * (1 + 1e-15) and 5.0 * (1 - 1e-15) <= lo <= 5.0
    # diag(2, 3) as CSR: sqrt(||A||_1 ||A||_inf) = 3, largest entry 3;
    # one entry per row and column -> 2 * 1 + 2 steps each way
    op = O.CsrOperator(2, 2, np.array([0, 1, 2]), np.array([0, 1]), np.array([2.0, 3.0]))
    up, lo, sa, st = op.validation_error_model()
    assert (sa, st) == (4, 4)
    assert 3.0 <= up <= 3.0 * (1 + 1e-15) and 3.0 * (1 - 1e-15) <= lo <= 3.0


def _counted(op):
    return V.Counted(op, R.RunStatistics(), CL.SyntheticClock())


class _NoRefinement(O.DenseOperator):
    def supports_validation_refinement(self):
        return False


def _exact_ls_problem():
    A = np.array([[1, 2, 0], [0, 1, 1], [3, 0, 1], [1, 1, 1], [2, -1, 0], [0, 0, 4]], dtype=np.float64)
    b = np.array([1, -2, 3, 0, 5, 1], dtype=np.float64)
    Af = [[Fraction(int(v)) for v in row] for row in A]
    bf = [Fraction(int(v)) for v in b]
    # normal equations solved exactly (Gauss-Jordan on rationals)
    n = 3
    N = [[sum(Af[k][i] * Af[k][j] for k in range(6)) for j in range(n)] for i in range(n)]
    rhs = [sum(Af[k][i] * bf[k] for k in range(6)) for i in range(n)]
    M = [N[i] + [rhs[i]] for i in range(n)]
    for c in range(n):
        piv = M[c][c]
        M[c] = [v / piv for v in M[c]]
        for r in range(n):
            if r != c:
                f = M[r][c]
                M[r] = [vr - f * vc for vr, vc in zip(M[r], M[c])]
    x_exact = [M[i][n] for i in range(n)]
    return A, b, Af, bf, x_exact


def test_long_double_refinement_branch_runs_and_tightens():
    """x = the rounded exact least-squares solution: the FP64 gradient is
    rounding noise below its error radius, so the validator takes the
    long-double refinement branch (two more products) and the bound it
    certifies is far smaller than the FP64-only bound -- and still encloses
    the exact normal residual of x."""
    A, b, Af, bf, x_exact = _exact_ls_problem()
    x = np.array([float(v) for v in x_exact])
    c_ref = _counted(O.DenseOperator(A))
    res = V.validate_original_candidate(c_ref, b, x, 0.0, 1e-30)
    c_plain = _counted(_NoRefinement(A))
    base = V.validate_original_candidate(c_plain, b, x, 0.0, 1e-30)
    assert c_plain.stats.base_validation_a_columns == 1 and c_plain.stats.base_validation_at_columns == 1
    assert c_ref.stats.base_validation_a_columns == 2 and c_ref.stats.base_validation_at_columns == 2
    assert res.status != R.SUCCESS and base.status != R.SUCCESS          # 1e-30 is out of reach
    # long-double products shrink the radii by about 2^-64 / 2^-53 = 1/2048; the
    # binary64 casts of Ax, r and g keep the refined bound above that ratio
    assert res.least_squares_backward_error_upper_bound < 1e-2 * base.least_squares_backward_error_upper_bound
    # exact normal residual of the stored x
    xf = [Fraction(v) for v in x]
    r = [sum(Af[k][j] * xf[j] for j in range(3)) - bf[k] for k in range(6)]
    g = [sum(Af[k][i] * r[k] for k in range(6)) for i in range(3)]
    g2 = sum(v * v for v in g)
    r2 = sum(v * v for v in r)
    sigma_max = float(np.linalg.svd(A, compute_uv=False)[0])
    # LS backward bound = ||g||+ / (||A||- ||r||-) >= ||g|| / (sigma_max (1 + 1e-12) ||r||)
    true_ratio = math.sqrt(float(g2)) / (sigma_max * (1 + 1e-12) * math.sqrt(float(r2)))
    assert res.least_squares_backward_error_upper_bound >= true_ratio


def test_ridge_one_step_correction_improves_a_single_mode_error():
    """A = diag(10, 1), lambda = 1e-2, x = x* + 1e-6 e1: the error lies on an
    eigenvector of H = A^T A + lambda I (eigenvalue 100.01), so one exact
    line step along g removes it.  The base bound ||g|| / sqrt(lambda) =
    100.01e-6 / 0.1 ~ 1e-3; the corrected bound is the energy error itself,
    sqrt(100.01) * 1e-6 ~ 1.00005e-5, up to the rounding radii."""
    lam = 1e-2
    A = np.diag([10.0, 1.0])
    b = np.array([1.0, 1.0])
    x_star = [Fraction(10) / Fraction(10001, 100), Fraction(1) / Fraction(101, 100)]
    x = np.array([float(x_star[0]) + 1e-6, float(x_star[1])])
    res = V.validate_original_candidate(_counted(O.DenseOperator(A)), b, x, lam, 1e-12)
    assert res.ridge_correction_disposition == R.RC_ATTEMPTED_IMPROVED
    e = [Fraction(x[0]) - x_star[0], Fraction(x[1]) - x_star[1]]
    energy = math.sqrt(float((Fraction(100) + Fraction(lam)) * e[0] ** 2 + (Fraction(1) + Fraction(lam)) * e[1] ** 2))
    assert res.ridge_corrected_energy_error_upper_bound >= energy
    assert res.ridge_corrected_energy_error_upper_bound <= energy * (1 + 1e-6)
    assert res.ridge_base_energy_error_upper_bound >= 50 * res.ridge_corrected_energy_error_upper_bound
    assert res.energy_error_upper_bound == res.ridge_corrected_energy_error_upper_bound
    assert res.ridge_correction_gamma == pytest.approx(1 / 100.01, rel=1e-9)
