"""Certified FP64 candidate validation (oracle).

Follows src/algorithms/mathematics/krylov/candidate_validation.cpp:
validate_original_candidate (:1176-1378), populate_candidate_norms
(:1151-1172), evaluate_accuracy_contract (:594-692),
build_base_validation_envelope (:822-926),
evaluate_certified_accuracy_contract (:694-814), refine_base_validation
(:951-1149) and one_step_ridge_energy_bound (:395-566).

fma(lambda, x, A^T r) is evaluated in numpy longdouble and rounded once to
binary64; where the 64-bit product is inexact this can differ from a true
fused multiply-add by one ulp, which the tests' tolerances absorb.

Pinned by tests/test_oracle_solvers.py: the certified bounds enclose the
true least-squares and compatible backward errors and the true ridge energy
error computed independently (numpy 2-norm, closed-form ridge solution) on
seeded problems, and an exact solution of a diagonal system certifies the
compatible backward error below 1e-14 with a zero normal residual
(test_validation_bounds_enclose_true_backward_error,
test_ridge_energy_bound_encloses_truth); by tests/test_oracle_pins.py: the
callback and one-operation radii against exact rationals, the dense and CSR
operator models on hand-computed matrices, a case that forces the
long-double refinement (exact-rational enclosure of the normal residual)
and one that forces the one-step ridge correction (single-mode error, the
corrected bound equal to the exact energy error up to the radii).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import fpbounds as fp
from . import results as R

INF = math.inf
NAN = math.nan
LD = np.longdouble


def validate_common_inputs(op, b, regularization, tolerance):
    """candidate_validation.cpp:570-592."""
    if op.rows <= 0 or op.cols <= 0:
        raise ValueError("operator dimensions must be positive")
    if b is None:
        raise ValueError("right-hand side cannot be null")
    if not math.isfinite(regularization) or regularization < 0.0:
        raise ValueError("regularization must be finite and nonnegative")
    if not math.isfinite(tolerance) or tolerance <= 0.0 or tolerance >= 1.0:
        raise ValueError("tolerance must lie strictly between zero and one")
    if not np.all(np.isfinite(b)):
        raise ValueError("right-hand side must be finite")


def model_available(model) -> bool:
    up, lo, sa, st = model
    return (math.isfinite(up) and up >= 0.0 and math.isfinite(lo) and lo >= 0.0 and lo <= up
            and sa > 0 and st > 0)


def underflow_radius(operations: int, output_size: int) -> float:
    if operations <= 0 or output_size <= 0:
        return 0.0
    return fp.upward_from_long_double(LD(operations) * np.sqrt(LD(output_size)) * LD(fp.DENORM_MIN))


def callback_error_radius(model, transpose: bool, input_norm_upper: float, output_size: int):
    if not model_available(model) or not math.isfinite(input_norm_upper) or input_norm_upper < 0.0:
        return INF
    if model[0] == 0.0 or input_norm_upper == 0.0:
        return 0.0
    ops = model[3] if transpose else model[2]
    gamma = fp.binary64_gamma(ops)
    main = fp.upward_multiply(gamma, fp.upward_multiply(model[0], input_norm_upper))
    return fp.upward_add(main, underflow_radius(ops, output_size))


def one_operation_radius(stored_upper: float, can_underflow: bool, output_size: int) -> float:
    if not math.isfinite(stored_upper) or stored_upper < 0.0:
        return INF
    r = fp.upward_multiply(fp.binary64_gamma(2), stored_upper)
    if can_underflow:
        r = fp.upward_add(r, underflow_radius(1, output_size))
    return r


def enlarge(stored, radius):
    if not math.isfinite(radius) or radius < 0.0:
        return (0.0, INF)
    return (fp.downward_difference(stored[0], radius), fp.upward_add(stored[1], radius))


def augmented_norm_interval(primary, secondary, lam):
    lo2 = fp.downward_add(fp.downward_multiply(primary[0], primary[0]),
                          fp.downward_multiply(lam, fp.downward_multiply(secondary[0], secondary[0])))
    up2 = fp.upward_add(fp.upward_multiply(primary[1], primary[1]),
                        fp.upward_multiply(lam, fp.upward_multiply(secondary[1], secondary[1])))
    return (fp.downward_sqrt(lo2), fp.upward_sqrt(up2))


def certified_ratio(num, den):
    if not math.isfinite(num):
        return INF
    if num == 0.0:
        return 0.0
    if not math.isfinite(den) or den <= 0.0:
        return INF
    return num / den


def relative_energy_bound(absolute, energy):
    if not math.isfinite(energy):
        return INF
    if absolute == 0.0:
        return 0.0
    if not math.isfinite(absolute) or energy <= 0.0:
        return INF
    s = absolute / energy
    if not math.isfinite(s) or s >= 1.0:
        return INF
    return s / (1.0 - s)


def long_double_gamma(operations: int) -> float:
    if operations < 0:
        return INF
    p = LD(operations) * (LD(0.5) * fp.LD_EPS)
    if not (p < 1):
        return INF
    return fp.upward_from_long_double(p / (LD(1) - p))


def long_double_underflow_radius(operations, output_size):
    if operations <= 0 or output_size <= 0:
        return 0.0
    tiny = LD(np.finfo(np.longdouble).smallest_subnormal)
    return fp.upward_from_long_double(LD(operations) * np.sqrt(LD(output_size)) * tiny)


def long_double_callback_error_radius(model, dot_length, input_norm_upper, output_size):
    if not model_available(model) or dot_length <= 0 or not math.isfinite(input_norm_upper) \
            or input_norm_upper < 0.0:
        return INF
    if model[0] == 0.0 or input_norm_upper == 0.0:
        return 0.0
    ops = 2 * dot_length
    main = fp.upward_multiply(long_double_gamma(ops), fp.upward_multiply(model[0], input_norm_upper))
    return fp.upward_add(main, long_double_underflow_radius(ops, output_size))


def long_double_one_operation_radius(stored_upper, can_underflow, output_size):
    if not math.isfinite(stored_upper) or stored_upper < 0.0:
        return INF
    r = fp.upward_multiply(long_double_gamma(2), stored_upper)
    if can_underflow:
        r = fp.upward_add(r, long_double_underflow_radius(1, output_size))
    return r


def fma_vec(a: float, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    return (LD(a) * x.astype(LD) + y.astype(LD)).astype(np.float64)


@dataclass
class Envelope:
    certified: bool = False
    operator_norm_upper_bound: float = INF
    operator_norm_lower_bound: float = 0.0
    residual_lower: float = 0.0
    residual_upper: float = INF
    gradient_upper: float = INF
    solution_lower: float = 0.0
    solution_upper: float = INF
    right_hand_side_lower: float = 0.0
    applied_solution_lower: float = 0.0
    applied_solution_upper: float = INF
    solution_energy_lower: float = 0.0
    gradient_error_radius: float = INF
    model: tuple = field(default_factory=lambda: (INF, 0.0, 0, 0))


class Counted:
    """Operator-call accounting shared by the search and validation paths."""

    def __init__(self, op, stats: R.RunStatistics, clock):
        self.op, self.stats, self.clock = op, stats, clock

    def validation_apply(self, x, transpose: bool, phase: str):
        s, clk = self.stats, self.clock
        start = clk.now()
        clk.charge_operator(self.op.relative_block_product_cost(1))
        y = (self.op.apply_transpose if transpose else self.op.apply)(x.reshape(-1, 1)).ravel()
        sec = clk.now() - start
        self._count(transpose, phase, sec)
        return y

    def validation_refine(self, x, transpose: bool):
        start = self.clock.now()
        self.clock.charge_operator(self.op.relative_block_product_cost(1))
        fn = self.op.apply_transpose_validation_refinement if transpose else \
            self.op.apply_validation_refinement
        y = fn(x)
        self._count(transpose, "base", self.clock.now() - start)
        return y

    def _count(self, transpose, phase, sec):
        s = self.stats
        if transpose:
            s.at_seconds += sec
            s.at_columns += 1
            s.validation_at_columns += 1
            s.at_block_calls += 1
            s.validation_at_block_calls += 1
            if phase == "base":
                s.base_validation_at_columns += 1
                s.base_validation_at_block_calls += 1
                s.base_validation_seconds += sec
            else:
                s.ridge_correction_at_columns += 1
                s.ridge_correction_at_block_calls += 1
                s.ridge_correction_validation_seconds += sec
        else:
            s.a_seconds += sec
            s.a_columns += 1
            s.validation_a_columns += 1
            s.a_block_calls += 1
            s.validation_a_block_calls += 1
            if phase == "base":
                s.base_validation_a_columns += 1
                s.base_validation_a_block_calls += 1
                s.base_validation_seconds += sec
            else:
                s.ridge_correction_a_columns += 1
                s.ridge_correction_a_block_calls += 1
                s.ridge_correction_validation_seconds += sec
        s.validation_seconds = s.base_validation_seconds + s.ridge_correction_validation_seconds


def evaluate_accuracy_contract(res: R.LeastSquaresResult, bnorm, lam, tol, ridge_bound=None):
    """candidate_validation.cpp:594-692 (nominal, uncertified contract)."""
    root = math.sqrt(lam)
    res.augmented_residual_norm = math.hypot(res.residual_norm, root * res.solution_norm)
    res.augmented_operator_norm_lower_bound = math.hypot(res.operator_norm_lower_bound, root)
    res.compatible_backward_error_upper_bound = certified_ratio(
        res.augmented_residual_norm,
        res.augmented_operator_norm_lower_bound * res.solution_norm + bnorm)
    res.least_squares_backward_error_upper_bound = certified_ratio(
        res.normal_residual_norm, res.augmented_operator_norm_lower_bound * res.augmented_residual_norm)
    res.backward_error_upper_bound = min(res.compatible_backward_error_upper_bound,
                                         res.least_squares_backward_error_upper_bound)
    den = (res.operator_norm_lower_bound ** 2 + lam) * res.solution_norm + \
        res.operator_norm_lower_bound * bnorm
    res.relative_normal_residual_upper_bound = certified_ratio(res.normal_residual_norm, den)
    if lam > 0.0:
        req = ridge_bound if ridge_bound is not None else res.normal_residual_norm / root
        res.energy_error_upper_bound = req if req >= 0.0 else INF
        res.objective_gap_upper_bound = res.energy_error_upper_bound ** 2
        res.relative_energy_error_upper_bound = relative_energy_bound(
            res.energy_error_upper_bound, res.solution_energy_norm)
    elif res.normal_residual_norm == 0.0:
        res.energy_error_upper_bound = 0.0
        res.relative_energy_error_upper_bound = 0.0
        res.objective_gap_upper_bound = 0.0
    else:
        res.energy_error_upper_bound = INF
        res.relative_energy_error_upper_bound = INF
        res.objective_gap_upper_bound = INF

    def fin(v):
        return math.isfinite(v) and v >= 0.0
    prims = [bnorm, res.residual_norm, res.augmented_residual_norm, res.normal_residual_norm,
             res.solution_norm, res.solution_energy_norm, res.objective,
             res.operator_norm_lower_bound, res.augmented_operator_norm_lower_bound]
    if not all(fin(v) for v in prims):
        res.status = R.NUMERICAL_BREAKDOWN
        res.stop_reason = R.STOP_NUMERICAL_BREAKDOWN
        return
    passed = (res.relative_energy_error_upper_bound <= tol) if lam > 0.0 else \
        (res.backward_error_upper_bound <= tol)
    res.status = R.SUCCESS if passed else R.WORK_LIMIT
    if res.normal_residual_norm == 0.0:
        res.stop_reason = R.EXACT_STATIONARITY
    elif res.status == R.SUCCESS and lam > 0.0:
        res.stop_reason = R.RELATIVE_ENERGY_ERROR
    elif res.status == R.SUCCESS and res.compatible_backward_error_upper_bound <= \
            res.least_squares_backward_error_upper_bound:
        res.stop_reason = R.COMPATIBLE_BACKWARD_ERROR
    elif res.status == R.SUCCESS:
        res.stop_reason = R.LEAST_SQUARES_BACKWARD_ERROR
    else:
        res.stop_reason = R.MAXIMUM_DEPTH


def evaluate_certified_accuracy_contract(res, env: Envelope, lam, tol, ridge_bound=None):
    """candidate_validation.cpp:694-814."""
    if not env.certified:
        for f in ("compatible_backward_error_upper_bound", "least_squares_backward_error_upper_bound",
                  "backward_error_upper_bound", "relative_normal_residual_upper_bound",
                  "energy_error_upper_bound", "relative_energy_error_upper_bound",
                  "objective_gap_upper_bound"):
            setattr(res, f, INF)
        res.status = R.WORK_LIMIT
        res.stop_reason = R.MAXIMUM_DEPTH
        return
    solution = (env.solution_lower, env.solution_upper)
    residual = (env.residual_lower, env.residual_upper)
    aug_res = augmented_norm_interval(residual, solution, lam)
    lo = res.operator_norm_lower_bound
    res.augmented_operator_norm_lower_bound = fp.downward_sqrt(
        fp.downward_add(fp.downward_multiply(lo, lo), lam))
    comp_den = fp.downward_add(fp.downward_multiply(res.augmented_operator_norm_lower_bound,
                                                    env.solution_lower), env.right_hand_side_lower)
    res.compatible_backward_error_upper_bound = fp.upward_divide(aug_res[1], comp_den)
    ls_den = fp.downward_multiply(res.augmented_operator_norm_lower_bound, aug_res[0])
    res.least_squares_backward_error_upper_bound = fp.upward_divide(env.gradient_upper, ls_den)
    res.backward_error_upper_bound = min(res.compatible_backward_error_upper_bound,
                                         res.least_squares_backward_error_upper_bound)
    normal_den = fp.downward_add(
        fp.downward_multiply(fp.downward_add(fp.downward_multiply(lo, lo), lam), env.solution_lower),
        fp.downward_multiply(lo, env.right_hand_side_lower))
    res.relative_normal_residual_upper_bound = fp.upward_divide(env.gradient_upper, normal_den)
    if lam > 0.0:
        root_lo = fp.downward_sqrt(lam)
        base = fp.upward_divide(env.gradient_upper, root_lo)
        res.energy_error_upper_bound = ridge_bound if ridge_bound is not None else base
        res.objective_gap_upper_bound = fp.upward_multiply(res.energy_error_upper_bound,
                                                           res.energy_error_upper_bound)
        rel_den = fp.downward_difference(env.solution_energy_lower, res.energy_error_upper_bound)
        res.relative_energy_error_upper_bound = fp.upward_divide(res.energy_error_upper_bound, rel_den)
    elif env.gradient_upper == 0.0:
        res.energy_error_upper_bound = 0.0
        res.relative_energy_error_upper_bound = 0.0
        res.objective_gap_upper_bound = 0.0
    else:
        res.energy_error_upper_bound = INF
        res.relative_energy_error_upper_bound = INF
        res.objective_gap_upper_bound = INF
    passed = env.gradient_upper == 0.0 or (
        res.relative_energy_error_upper_bound <= tol if lam > 0.0 else
        res.backward_error_upper_bound <= tol)
    res.status = R.SUCCESS if passed else R.WORK_LIMIT
    if env.gradient_upper == 0.0:
        res.stop_reason = R.EXACT_STATIONARITY
    elif passed and lam > 0.0:
        res.stop_reason = R.RELATIVE_ENERGY_ERROR
    elif passed and res.compatible_backward_error_upper_bound <= \
            res.least_squares_backward_error_upper_bound:
        res.stop_reason = R.COMPATIBLE_BACKWARD_ERROR
    elif passed:
        res.stop_reason = R.LEAST_SQUARES_BACKWARD_ERROR
    else:
        res.stop_reason = R.MAXIMUM_DEPTH


def failed_numerically(res) -> bool:
    return res.status == R.NUMERICAL_BREAKDOWN or res.stop_reason == R.STOP_NUMERICAL_BREAKDOWN


def build_base_envelope(op, b, x, ax, r, atr, g, lam) -> Envelope:
    """candidate_validation.cpp:822-926."""
    env = Envelope()
    env.model = op.validation_error_model()
    env.operator_norm_upper_bound = env.model[0]
    if not model_available(env.model):
        return env
    m, n = op.rows, op.cols
    b_n, x_n, ax_n = fp.norm_interval(b), fp.norm_interval(x), fp.norm_interval(ax)
    r_n, atr_n, g_n = fp.norm_interval(r), fp.norm_interval(atr), fp.norm_interval(g)
    apply_err = callback_error_radius(env.model, False, x_n[1], m)
    sub_err = one_operation_radius(r_n[1], ax_n[1] > 0.0 or b_n[1] > 0.0, m)
    res_err = fp.upward_add(apply_err, sub_err)
    true_res = enlarge(r_n, res_err)
    t_err = callback_error_radius(env.model, True, r_n[1], n)
    ridge_err = one_operation_radius(g_n[1], lam > 0.0 and x_n[1] > 0.0, n)
    prop = fp.upward_multiply(env.operator_norm_upper_bound, res_err)
    g_err = fp.upward_add(fp.upward_add(t_err, prop), ridge_err)
    true_g = enlarge(g_n, g_err)
    true_ax = enlarge(ax_n, apply_err)
    op_lo = env.model[1]
    if x_n[1] > 0.0:
        op_lo = max(op_lo, fp.downward_divide(fp.downward_difference(ax_n[0], apply_err), x_n[1]))
    if r_n[1] > 0.0:
        op_lo = max(op_lo, fp.downward_divide(fp.downward_difference(atr_n[0], t_err), r_n[1]))
    energy = augmented_norm_interval(true_ax, x_n, lam)
    env.residual_lower, env.residual_upper = true_res
    env.gradient_upper = true_g[1]
    env.solution_lower, env.solution_upper = x_n
    env.right_hand_side_lower = b_n[0]
    env.applied_solution_lower, env.applied_solution_upper = true_ax
    env.solution_energy_lower = energy[0]
    env.gradient_error_radius = g_err
    env.certified = (math.isfinite(env.residual_upper) and math.isfinite(env.gradient_upper)
                     and math.isfinite(env.solution_upper)
                     and math.isfinite(env.operator_norm_upper_bound))
    env.applied_solution_lower = max(0.0, env.applied_solution_lower)
    env.right_hand_side_lower = max(0.0, env.right_hand_side_lower)
    env.operator_norm_lower_bound = op_lo
    return env


def refine_base_validation(counted: Counted, b, x, lam):
    """candidate_validation.cpp:951-1149 (long double refinement)."""
    op = counted.op
    if not op.supports_validation_refinement():
        return None
    model = op.validation_error_model()
    if not model_available(model):
        return None
    m, n = op.rows, op.cols
    b_n, x_n = fp.norm_interval(b), fp.norm_interval(x)
    if not math.isfinite(b_n[1]) or not math.isfinite(x_n[1]):
        return None
    lx = x.astype(LD)
    lax = counted.validation_refine(lx, False)
    if lax is None or not np.all(np.isfinite(lax)):
        return None
    ax = lax.astype(np.float64)
    if not np.all(np.isfinite(ax)):
        return None
    ax_n = fp.norm_interval(ax)
    ax_cast = one_operation_radius(ax_n[1], True, m)
    ax_acc = long_double_callback_error_radius(model, n, x_n[1], m)
    ax_err = fp.upward_add(ax_acc, ax_cast)
    lr = lax - b.astype(LD)
    r = lr.astype(np.float64)
    if not np.all(np.isfinite(lr)) or not np.all(np.isfinite(r)):
        return None
    r_n = fp.norm_interval(r)
    r_cast = one_operation_radius(r_n[1], True, m)
    lr_upper = fp.upward_add(r_n[1], r_cast)
    r_sub = long_double_one_operation_radius(lr_upper, ax_n[1] > 0.0 or b_n[1] > 0.0, m)
    lr_err = fp.upward_add(ax_acc, r_sub)
    r_err = fp.upward_add(lr_err, r_cast)
    lt = counted.validation_refine(lr, True)
    if lt is None or not np.all(np.isfinite(lt)):
        return None
    atr = lt.astype(np.float64)
    if not np.all(np.isfinite(atr)):
        return None
    t_n = fp.norm_interval(atr)
    t_cast = one_operation_radius(t_n[1], True, n)
    t_acc = long_double_callback_error_radius(model, m, lr_upper, n)
    lg = LD(lam) * lx + lt
    g = lg.astype(np.float64)
    if not np.all(np.isfinite(lg)) or not np.all(np.isfinite(g)):
        return None
    g_n = fp.norm_interval(g)
    g_cast = one_operation_radius(g_n[1], True, n)
    lg_upper = fp.upward_add(g_n[1], g_cast)
    ridge_err = long_double_one_operation_radius(lg_upper, lam > 0.0 and x_n[1] > 0.0, n)
    prop = fp.upward_multiply(model[0], lr_err)
    g_err = fp.upward_add(fp.upward_add(t_acc, prop), fp.upward_add(ridge_err, g_cast))
    true_ax = enlarge(ax_n, ax_err)
    true_r = enlarge(r_n, r_err)
    true_g = enlarge(g_n, g_err)
    op_lo = model[1]
    if x_n[1] > 0.0:
        op_lo = max(op_lo, fp.downward_divide(fp.downward_difference(ax_n[0], ax_err), x_n[1]))
    if r_n[1] > 0.0:
        conv = fp.upward_multiply(model[0], r_cast)
        te = fp.upward_add(fp.upward_add(t_acc, t_cast), conv)
        op_lo = max(op_lo, fp.downward_divide(fp.downward_difference(t_n[0], te), r_n[1]))
    energy = augmented_norm_interval(true_ax, x_n, lam)
    env = Envelope(model=model, operator_norm_upper_bound=model[0], operator_norm_lower_bound=op_lo,
                   residual_lower=true_r[0], residual_upper=true_r[1], gradient_upper=true_g[1],
                   solution_lower=x_n[0], solution_upper=x_n[1],
                   right_hand_side_lower=max(0.0, b_n[0]),
                   applied_solution_lower=max(0.0, true_ax[0]), applied_solution_upper=true_ax[1],
                   solution_energy_lower=energy[0], gradient_error_radius=g_err)
    env.certified = (math.isfinite(env.residual_upper) and math.isfinite(env.gradient_upper)
                     and math.isfinite(env.solution_upper)
                     and math.isfinite(env.operator_norm_upper_bound))
    if not env.certified:
        return None
    return env, ax, r, atr, g


def one_step_ridge_energy_bound(counted: Counted, g, gnorm, lam, env: Envelope):
    """candidate_validation.cpp:395-566; returns (disposition, bound, gamma, zhz, two_cross, qsq)."""
    out = dict(disposition=R.RC_ATTEMPTED_NONFINITE, bound=NAN, gamma=NAN, zhz=NAN, cross=NAN, qsq=NAN)
    op = counted.op
    m, n = op.rows, op.cols
    if not (gnorm > 0.0) or not math.isfinite(gnorm) or not np.all(np.isfinite(g)):
        out["disposition"] = R.RC_SKIPPED_NONFINITE_CANDIDATE
        return out
    d = g / gnorm
    if not np.all(np.isfinite(d)):
        out["disposition"] = R.RC_SKIPPED_NONFINITE_CANDIDATE
        return out
    d_n = fp.norm_interval(d)
    ad = counted.validation_apply(d, False, "ridge")
    if not np.all(np.isfinite(ad)):
        out["disposition"] = R.RC_ATTEMPTED_PARTIAL_FAILURE
        return out
    ad_n = fp.norm_interval(ad)
    ad_err = callback_error_radius(env.model, False, d_n[1], m)
    td = counted.validation_apply(ad, True, "ridge")
    t_err = callback_error_radius(env.model, True, ad_n[1], n)
    hd = fma_vec(lam, d, td)
    if not np.all(np.isfinite(hd)):
        return out
    hd_n = fp.norm_interval(hd)
    prop = fp.upward_multiply(env.operator_norm_upper_bound, ad_err)
    hsum = one_operation_radius(hd_n[1], lam > 0.0 and d_n[1] > 0.0, n)
    h_err = fp.upward_add(fp.upward_add(t_err, prop), hsum)
    curvature = float(np.dot(d, hd))
    if not math.isfinite(curvature):
        return out
    if not (curvature > 0.0):
        out["disposition"] = R.RC_ATTEMPTED_INVALID_CURVATURE
        return out
    step = gnorm / curvature
    if not (step > 0.0) or not math.isfinite(step):
        return out
    q = fma_vec(-step, hd, g)
    if not np.all(np.isfinite(q)):
        return out
    q_n = fp.norm_interval(q)
    axpy_err = one_operation_radius(q_n[1], gnorm > 0.0 and hd_n[1] > 0.0, n)
    q_err = fp.upward_add(fp.upward_add(env.gradient_error_radius, fp.upward_multiply(step, h_err)),
                          axpy_err)
    q_upper = fp.upward_add(q_n[1], q_err)
    ad_true = fp.upward_add(ad_n[1], ad_err)
    step2 = fp.upward_multiply(step, step)
    de2 = fp.upward_add(fp.upward_multiply(ad_true, ad_true),
                        fp.upward_multiply(lam, fp.upward_multiply(d_n[1], d_n[1])))
    zhz = fp.upward_multiply(step2, de2)
    cross = fp.upward_multiply(step, fp.upward_multiply(d_n[1], q_upper))
    two_cross = fp.upward_multiply(2.0, cross)
    qsq = fp.upward_multiply(q_upper, q_upper)
    qe = fp.upward_divide(qsq, lam)
    corrected = fp.upward_sqrt(fp.upward_add(zhz, fp.upward_add(two_cross, qe)))
    if math.isnan(corrected):
        return out
    out.update(disposition=R.RC_ATTEMPTED_NOT_IMPROVED, bound=corrected, gamma=1.0 / curvature,
               zhz=zhz, cross=two_cross, qsq=qsq)
    return out


def populate_norms(res, x, ax, r, g, lam):
    axn = float(np.linalg.norm(ax))
    res.residual_norm = float(np.linalg.norm(r))
    res.normal_residual_norm = float(np.linalg.norm(g))
    res.solution_norm = float(np.linalg.norm(x))
    res.solution_energy_norm = math.hypot(axn, math.sqrt(lam) * res.solution_norm)
    res.objective = res.residual_norm * res.residual_norm + lam * res.solution_norm * res.solution_norm


def validate_original_candidate(counted: Counted, b, x, lam, tol) -> R.LeastSquaresResult:
    """candidate_validation.cpp:1176-1378."""
    op = counted.op
    x = np.asarray(x, dtype=np.float64).ravel()
    if x.size != op.cols:
        raise ValueError("candidate solution has the wrong size")
    res = R.LeastSquaresResult(rows=op.rows, cols=op.cols, solution=x.copy())
    res.operator_norm_lower_bound = 0.0
    ax = counted.validation_apply(x, False, "base")
    r = ax - b
    atr = counted.validation_apply(r, True, "base")
    g = fma_vec(lam, x, atr)
    populate_norms(res, x, ax, r, g, lam)
    bnorm = float(np.linalg.norm(b))
    evaluate_accuracy_contract(res, bnorm, lam, tol)
    finite = all(np.all(np.isfinite(v)) for v in (x, ax, r, atr, g))
    if not finite or failed_numerically(res):
        res.status = R.NUMERICAL_BREAKDOWN
        res.stop_reason = R.STOP_NUMERICAL_BREAKDOWN
    else:
        env = build_base_envelope(op, b, x, ax, r, atr, g, lam)
        res.operator_norm_lower_bound = env.operator_norm_lower_bound
        evaluate_certified_accuracy_contract(res, env, lam, tol)
        if res.status != R.SUCCESS and env.certified and \
                env.gradient_error_radius >= res.normal_residual_norm:
            refined = refine_base_validation(counted, b, x, lam)
            if refined is not None:
                env, ax, r, atr, g = refined
                populate_norms(res, x, ax, r, g, lam)
                evaluate_accuracy_contract(res, bnorm, lam, tol)
                res.operator_norm_lower_bound = env.operator_norm_lower_bound
                evaluate_certified_accuracy_contract(res, env, lam, tol)
        if lam > 0.0:
            res.ridge_base_energy_error_upper_bound = res.energy_error_upper_bound
            if res.status == R.SUCCESS:
                res.ridge_correction_disposition = R.RC_SKIPPED_BASE_PASS
            elif not env.certified:
                res.ridge_correction_disposition = R.RC_SKIPPED_NONFINITE_CANDIDATE
            elif env.solution_energy_lower == 0.0:
                res.ridge_correction_disposition = R.RC_SKIPPED_ZERO_ENERGY
            elif not (env.gradient_upper > 0.0) or not math.isfinite(env.gradient_upper):
                res.ridge_correction_disposition = R.RC_SKIPPED_NONFINITE_CANDIDATE
            else:
                a = one_step_ridge_energy_bound(counted, g, res.normal_residual_norm, lam, env)
                res.ridge_correction_disposition = a["disposition"]
                res.ridge_corrected_energy_error_upper_bound = a["bound"]
                res.ridge_correction_gamma = a["gamma"]
                res.ridge_correction_z_h_z = a["zhz"]
                res.ridge_correction_two_abs_z_t_q = a["cross"]
                res.ridge_correction_q_norm_squared = a["qsq"]
                if a["disposition"] == R.RC_ATTEMPTED_NOT_IMPROVED and math.isfinite(a["bound"]) \
                        and a["bound"] < res.energy_error_upper_bound:
                    res.ridge_correction_disposition = R.RC_ATTEMPTED_IMPROVED
                    evaluate_certified_accuracy_contract(res, env, lam, tol, a["bound"])
    if lam > 0.0 and res.status == R.NUMERICAL_BREAKDOWN:
        res.ridge_base_energy_error_upper_bound = NAN if math.isnan(res.normal_residual_norm) else INF
        res.energy_error_upper_bound = res.ridge_base_energy_error_upper_bound
        res.relative_energy_error_upper_bound = INF
        res.objective_gap_upper_bound = INF
        res.ridge_correction_disposition = R.RC_SKIPPED_NONFINITE_CANDIDATE
    res.statistics = R.RunStatistics(**vars(counted.stats))
    return res
