"""Result types and the Torch result-tensor layout (oracle).

Follows include/amfls/least_squares_result.hpp:8-229 and
bindings/torch/result_tensors.hpp:12-156 (metrics 21, counters 36,
timings 9, trace_float 29, trace_int 4).  Pure data; pinned by the widths and
field positions checked in tests/test_oracle_solvers.py
(test_result_tensor_layout_widths).
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

NAN = math.nan
INF = math.inf

# SolverStatus (least_squares_result.hpp:8-14)
SUCCESS, WORK_LIMIT, BASIS_LIMIT, PRECISION_LIMIT, NUMERICAL_BREAKDOWN = range(5)
# StopReason (least_squares_result.hpp:16-28)
(CERTIFIED_OPTIMALITY, EXACT_STATIONARITY, EXHAUSTED_SEARCH_SPACE, MAXIMUM_DEPTH,
 MAXIMUM_EPOCHS, MAXIMUM_BASIS, STOP_PRECISION_LIMIT, STOP_NUMERICAL_BREAKDOWN,
 COMPATIBLE_BACKWARD_ERROR, LEAST_SQUARES_BACKWARD_ERROR, RELATIVE_ENERGY_ERROR) = range(11)
# RidgeCorrectionDisposition (least_squares_result.hpp:31-55)
(RC_NOT_APPLICABLE, RC_SKIPPED_BASE_PASS, RC_SKIPPED_ZERO_ENERGY, RC_SKIPPED_NONFINITE_CANDIDATE,
 RC_ATTEMPTED_IMPROVED, RC_ATTEMPTED_NOT_IMPROVED, RC_ATTEMPTED_INVALID_CURVATURE,
 RC_ATTEMPTED_PARTIAL_FAILURE, RC_ATTEMPTED_NONFINITE) = range(9)

COUNTER_FIELDS = [
    "a_columns", "at_columns", "a_block_calls", "at_block_calls",
    "search_a_columns", "search_at_columns", "validation_a_columns", "validation_at_columns",
    "sketch_a_columns", "sketch_at_columns", "sketch_a_block_calls", "sketch_at_block_calls",
    "iterative_a_columns", "iterative_at_columns", "iterative_a_block_calls",
    "iterative_at_block_calls", "gaussian_random_columns", "gaussian_random_values",
    "search_a_block_calls", "search_at_block_calls", "validation_a_block_calls",
    "validation_at_block_calls", "base_validation_a_columns", "base_validation_at_columns",
    "base_validation_a_block_calls", "base_validation_at_block_calls",
    "ridge_correction_a_columns", "ridge_correction_at_columns",
    "ridge_correction_a_block_calls", "ridge_correction_at_block_calls",
    "gaussian_random_block_requests",
]
SECONDS_FIELDS = ["total_seconds", "a_seconds", "at_seconds", "orthogonalization_seconds",
                  "projected_solve_seconds", "validation_seconds", "base_validation_seconds",
                  "ridge_correction_validation_seconds"]


@dataclass
class RunStatistics:
    a_columns: int = 0
    at_columns: int = 0
    a_block_calls: int = 0
    at_block_calls: int = 0
    search_a_columns: int = 0
    search_at_columns: int = 0
    validation_a_columns: int = 0
    validation_at_columns: int = 0
    sketch_a_columns: int = 0
    sketch_at_columns: int = 0
    sketch_a_block_calls: int = 0
    sketch_at_block_calls: int = 0
    iterative_a_columns: int = 0
    iterative_at_columns: int = 0
    iterative_a_block_calls: int = 0
    iterative_at_block_calls: int = 0
    gaussian_random_columns: int = 0
    gaussian_random_values: int = 0
    total_seconds: float = 0.0
    a_seconds: float = 0.0
    at_seconds: float = 0.0
    orthogonalization_seconds: float = 0.0
    projected_solve_seconds: float = 0.0
    validation_seconds: float = 0.0
    search_a_block_calls: int = 0
    search_at_block_calls: int = 0
    validation_a_block_calls: int = 0
    validation_at_block_calls: int = 0
    base_validation_a_columns: int = 0
    base_validation_at_columns: int = 0
    base_validation_a_block_calls: int = 0
    base_validation_at_block_calls: int = 0
    ridge_correction_a_columns: int = 0
    ridge_correction_at_columns: int = 0
    ridge_correction_a_block_calls: int = 0
    ridge_correction_at_block_calls: int = 0
    base_validation_seconds: float = 0.0
    ridge_correction_validation_seconds: float = 0.0
    gaussian_random_block_requests: int = 0


@dataclass
class IterationRecord:
    epoch: int = 0
    depth: int = 0
    auxiliary_width: int = 0
    basis_rank: int = 0
    a_columns: int = 0
    at_columns: int = 0
    a_block_calls: int = 0
    at_block_calls: int = 0
    objective: float = 0.0
    residual_norm: float = 0.0
    augmented_residual_norm: float = 0.0
    normal_residual_norm: float = 0.0
    operator_norm_lower_bound: float = 0.0
    compatible_backward_error_upper_bound: float = 0.0
    least_squares_backward_error_upper_bound: float = 0.0
    backward_error_upper_bound: float = 0.0
    relative_energy_error_upper_bound: float = 0.0
    relative_normal_residual_upper_bound: float = 0.0
    solution_norm: float = 0.0
    solution_energy_norm: float = 0.0
    augmented_operator_norm_lower_bound: float = 0.0
    energy_error_upper_bound: float = 0.0
    objective_gap_upper_bound: float = 0.0
    contract_status: int = WORK_LIMIT
    contract_stop_reason: int = MAXIMUM_DEPTH
    contract_passed: bool = False
    ridge_correction_disposition: int = RC_NOT_APPLICABLE
    ridge_base_energy_error_upper_bound: float = NAN
    ridge_corrected_energy_error_upper_bound: float = NAN
    ridge_correction_gamma: float = NAN
    ridge_correction_z_h_z: float = NAN
    ridge_correction_two_abs_z_t_q: float = NAN
    ridge_correction_q_norm_squared: float = NAN


@dataclass
class LeastSquaresResult:
    rows: int = 0
    cols: int = 0
    iterations: int = 0
    depth: int = 0
    auxiliary_width: int = 0
    basis_rank: int = 0
    solution: np.ndarray = field(default_factory=lambda: np.zeros(0))
    objective: float = 0.0
    residual_norm: float = 0.0
    augmented_residual_norm: float = 0.0
    normal_residual_norm: float = 0.0
    solution_norm: float = 0.0
    solution_energy_norm: float = 0.0
    operator_norm_lower_bound: float = 0.0
    augmented_operator_norm_lower_bound: float = 0.0
    compatible_backward_error_upper_bound: float = 0.0
    least_squares_backward_error_upper_bound: float = 0.0
    backward_error_upper_bound: float = 0.0
    relative_normal_residual_upper_bound: float = 0.0
    energy_error_upper_bound: float = 0.0
    relative_energy_error_upper_bound: float = 0.0
    objective_gap_upper_bound: float = 0.0
    status: int = WORK_LIMIT
    stop_reason: int = MAXIMUM_DEPTH
    statistics: RunStatistics = field(default_factory=RunStatistics)
    trace: list = field(default_factory=list)
    ridge_correction_disposition: int = RC_NOT_APPLICABLE
    ridge_base_energy_error_upper_bound: float = NAN
    ridge_corrected_energy_error_upper_bound: float = NAN
    ridge_correction_gamma: float = NAN
    ridge_correction_z_h_z: float = NAN
    ridge_correction_two_abs_z_t_q: float = NAN
    ridge_correction_q_norm_squared: float = NAN

    def copy(self) -> "LeastSquaresResult":
        return copy.deepcopy(self)


def make_iteration_record(result: LeastSquaresResult, epoch: int, depth: int,
                          auxiliary_width: int) -> IterationRecord:
    """krylov_workspace.cpp:1206-1256."""
    s = result.statistics
    return IterationRecord(
        epoch=epoch, depth=depth, auxiliary_width=auxiliary_width, basis_rank=result.basis_rank,
        a_columns=s.a_columns, at_columns=s.at_columns, a_block_calls=s.a_block_calls,
        at_block_calls=s.at_block_calls, contract_status=result.status,
        contract_stop_reason=result.stop_reason, contract_passed=result.status == SUCCESS,
        ridge_correction_disposition=result.ridge_correction_disposition,
        objective=result.objective, residual_norm=result.residual_norm,
        augmented_residual_norm=result.augmented_residual_norm,
        normal_residual_norm=result.normal_residual_norm,
        operator_norm_lower_bound=result.operator_norm_lower_bound,
        compatible_backward_error_upper_bound=result.compatible_backward_error_upper_bound,
        least_squares_backward_error_upper_bound=result.least_squares_backward_error_upper_bound,
        backward_error_upper_bound=result.backward_error_upper_bound,
        relative_energy_error_upper_bound=result.relative_energy_error_upper_bound,
        relative_normal_residual_upper_bound=result.relative_normal_residual_upper_bound,
        solution_norm=result.solution_norm, solution_energy_norm=result.solution_energy_norm,
        augmented_operator_norm_lower_bound=result.augmented_operator_norm_lower_bound,
        energy_error_upper_bound=result.energy_error_upper_bound,
        objective_gap_upper_bound=result.objective_gap_upper_bound,
        ridge_base_energy_error_upper_bound=result.ridge_base_energy_error_upper_bound,
        ridge_corrected_energy_error_upper_bound=result.ridge_corrected_energy_error_upper_bound,
        ridge_correction_gamma=result.ridge_correction_gamma,
        ridge_correction_z_h_z=result.ridge_correction_z_h_z,
        ridge_correction_two_abs_z_t_q=result.ridge_correction_two_abs_z_t_q,
        ridge_correction_q_norm_squared=result.ridge_correction_q_norm_squared)


METRIC_FIELDS = [
    "objective", "residual_norm", "augmented_residual_norm", "normal_residual_norm",
    "solution_norm", "solution_energy_norm", "operator_norm_lower_bound",
    "augmented_operator_norm_lower_bound", "compatible_backward_error_upper_bound",
    "least_squares_backward_error_upper_bound", "backward_error_upper_bound",
    "relative_normal_residual_upper_bound", "energy_error_upper_bound",
    "relative_energy_error_upper_bound", "objective_gap_upper_bound",
    "ridge_base_energy_error_upper_bound", "ridge_corrected_energy_error_upper_bound",
    "ridge_correction_gamma", "ridge_correction_z_h_z", "ridge_correction_two_abs_z_t_q",
    "ridge_correction_q_norm_squared",
]

TRACE_FLOAT_FIELDS = [
    "epoch", "depth", "auxiliary_width", "basis_rank", "a_columns", "at_columns",
    "a_block_calls", "at_block_calls", "objective", "residual_norm", "augmented_residual_norm",
    "normal_residual_norm", "operator_norm_lower_bound", "compatible_backward_error_upper_bound",
    "least_squares_backward_error_upper_bound", "backward_error_upper_bound",
    "relative_normal_residual_upper_bound", "relative_energy_error_upper_bound",
    "solution_norm", "solution_energy_norm", "augmented_operator_norm_lower_bound",
    "energy_error_upper_bound", "objective_gap_upper_bound",
    "ridge_base_energy_error_upper_bound", "ridge_corrected_energy_error_upper_bound",
    "ridge_correction_gamma", "ridge_correction_z_h_z", "ridge_correction_two_abs_z_t_q",
    "ridge_correction_q_norm_squared",
]


def to_tensors(result: LeastSquaresResult):
    """result_tensors.hpp:24-156 as numpy arrays (8-tuple order)."""
    s = result.statistics
    metrics = np.array([getattr(result, f) for f in METRIC_FIELDS], dtype=np.float64)
    counters = np.array([
        s.a_columns, s.at_columns, s.a_block_calls, s.at_block_calls, s.search_a_columns,
        s.search_at_columns, s.validation_a_columns, s.validation_at_columns, result.iterations,
        result.depth, result.auxiliary_width, result.basis_rank, s.sketch_a_columns,
        s.sketch_at_columns, s.sketch_a_block_calls, s.sketch_at_block_calls,
        s.iterative_a_columns, s.iterative_at_columns, s.iterative_a_block_calls,
        s.iterative_at_block_calls, s.gaussian_random_columns, s.gaussian_random_values,
        s.search_a_block_calls, s.search_at_block_calls, s.validation_a_block_calls,
        s.validation_at_block_calls, s.base_validation_a_columns, s.base_validation_at_columns,
        s.base_validation_a_block_calls, s.base_validation_at_block_calls,
        s.ridge_correction_a_columns, s.ridge_correction_at_columns,
        s.ridge_correction_a_block_calls, s.ridge_correction_at_block_calls,
        s.gaussian_random_block_requests, result.ridge_correction_disposition,
    ], dtype=np.int64)
    other = max(0.0, s.total_seconds - s.a_seconds - s.at_seconds - s.orthogonalization_seconds
                - s.projected_solve_seconds)
    timings = np.array([s.total_seconds, s.a_seconds, s.at_seconds, s.orthogonalization_seconds,
                        s.projected_solve_seconds, s.validation_seconds, other,
                        s.base_validation_seconds, s.ridge_correction_validation_seconds])
    tf = np.array([[float(getattr(r, f)) for f in TRACE_FLOAT_FIELDS] for r in result.trace],
                  dtype=np.float64).reshape(len(result.trace), 29)
    ti = np.array([[r.contract_status, r.contract_stop_reason, int(r.contract_passed),
                    r.ridge_correction_disposition] for r in result.trace],
                  dtype=np.int64).reshape(len(result.trace), 4)
    return (result.solution.copy(), metrics, counters, timings, tf, ti, result.status,
            result.stop_reason)
