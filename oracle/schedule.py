"""Certificate checkpoint cadence (oracle).

Follows src/algorithms/mathematics/krylov/certificate_checkpoint_schedule.hpp:15-131.
Pinned by tests/test_oracle_controller.py: an independent enumeration of the
static cadence (powers of two, plus 3*2^j for lambda = 0) and hand-computed
log-linear crossings (test_static_checkpoints_enumeration,
test_predicted_crossing_log_linear).
"""
from __future__ import annotations

import math

import numpy as np

INT_MAX = 2**31 - 1


def is_certificate_checkpoint(iteration: int, regularization: float) -> bool:
    if iteration <= 1:
        return True
    odd = iteration
    while odd % 2 == 0:
        odd //= 2
    return odd == 1 or (regularization == 0.0 and odd == 3)


def predicted_crossing(previous: float, current: float, previous_iteration: int,
                       current_iteration: int, tolerance: float) -> int:
    if not (previous > current) or not (current > tolerance) or not (tolerance > 0.0) \
            or not math.isfinite(previous) or not math.isfinite(current) \
            or not math.isfinite(tolerance):
        return INT_MAX
    ld = np.longdouble
    log_distance = np.log(ld(current)) - np.log(ld(tolerance))
    log_contraction = np.log(ld(previous)) - np.log(ld(current))
    distance = ld(current_iteration - previous_iteration)
    offset = np.ceil(distance * log_distance / log_contraction)
    if not (offset > 0) or not np.isfinite(offset) or offset > ld(INT_MAX - current_iteration):
        return INT_MAX
    return current_iteration + int(offset)


class CertificateCheckpointSchedule:
    def __init__(self, regularization: float, tolerance: float):
        self.regularization = regularization
        self.tolerance = tolerance
        self.predicted_iteration = INT_MAX
        self.previous_iteration = 0
        self.previous_compatible = 0.0
        self.previous_least_squares = 0.0
        self.previous_relative_energy = 0.0
        self.has_previous = False

    def should_evaluate(self, iteration: int) -> bool:
        return is_certificate_checkpoint(iteration, self.regularization) or \
            iteration >= self.predicted_iteration

    def record_evaluation(self, iteration: int, result) -> None:
        self.predicted_iteration = INT_MAX
        static = is_certificate_checkpoint(iteration, self.regularization)
        if static and self.has_previous and iteration > self.previous_iteration:
            if self.regularization > 0.0:
                self.predicted_iteration = predicted_crossing(
                    self.previous_relative_energy, result.relative_energy_error_upper_bound,
                    self.previous_iteration, iteration, self.tolerance)
            else:
                a = predicted_crossing(
                    self.previous_compatible, result.compatible_backward_error_upper_bound,
                    self.previous_iteration, iteration, self.tolerance)
                b = predicted_crossing(
                    self.previous_least_squares, result.least_squares_backward_error_upper_bound,
                    self.previous_iteration, iteration, self.tolerance)
                self.predicted_iteration = min(a, b)
        self.previous_iteration = iteration
        self.previous_compatible = result.compatible_backward_error_upper_bound
        self.previous_least_squares = result.least_squares_backward_error_upper_bound
        self.previous_relative_energy = result.relative_energy_error_upper_bound
        self.has_previous = True
