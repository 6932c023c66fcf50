"""Outward-rounded FP64 interval helpers (oracle).

Follows src/algorithms/mathematics/floating_point_bounds.hpp:12-403.  The
reference evaluates every helper in x87 ``long double`` and then steps one
binary64 ulp outward with ``nextafter``; numpy's ``longdouble`` is the same
80-bit x87 format on x86-64 Linux, so the arithmetic is reproduced exactly.

Pinned by tests/test_oracle_fpbounds.py: lower <= exact <= upper against
exact rational arithmetic (``fractions.Fraction``) on random and adversarial
inputs, the special cases 0 / inf / subnormal named in the reference, and
gamma_k against its closed form k*u/(1-k*u).

One deliberate difference: ``norm_interval`` accumulates the scaled square sum
with ``numpy.dot`` instead of a sequential FMA chain.  Any summation order of k
rounded products obeys the same |error| <= gamma_k * exact bound, so the
reference's positive_sum_{lower,upper}_bound(sum, k) formulas stay valid.
"""
from __future__ import annotations

import math

import numpy as np

LD = np.longdouble
INF = math.inf
DBL_MAX = float(np.finfo(np.float64).max)
DENORM_MIN = 5e-324
LD_EPS = LD(np.finfo(np.longdouble).eps)  # 2**-63
U64 = LD(2.0) ** -53  # binary64 unit roundoff, 0.5 * DBL_EPSILON


def _ld(x) -> np.longdouble:
    return LD(x)


def _isfinite_ld(v) -> bool:
    return bool(np.isfinite(v))


def upward_from_long_double(value) -> float:
    """floating_point_bounds.hpp:35-51."""
    value = _ld(value)
    if not (value >= 0) or not _isfinite_ld(value) or value > LD(DBL_MAX):
        return INF
    if value == 0:
        return 0.0
    rounded = float(value)
    if LD(rounded) < value:
        rounded = float(np.nextafter(rounded, INF))
    return float(np.nextafter(rounded, INF))


def downward_from_long_double(value) -> float:
    """floating_point_bounds.hpp:53-66."""
    value = _ld(value)
    if not (value > 0) or not _isfinite_ld(value):
        return 0.0
    if value > LD(DBL_MAX):
        return DBL_MAX
    rounded = float(value)
    if LD(rounded) > value:
        rounded = float(np.nextafter(rounded, 0.0))
    return float(np.nextafter(rounded, 0.0))


def accumulation_gamma(operations: int) -> np.longdouble:
    """floating_point_bounds.hpp:68-77 (long-double unit roundoff)."""
    unit = LD(0.5) * LD_EPS
    product = LD(operations) * unit
    if not (product < 1):
        return LD(INF)
    return product / (LD(1) - product)


def binary64_gamma(operations: int) -> float:
    """floating_point_bounds.hpp:123-135: gamma_k = k u / (1 - k u), rounded up."""
    if operations < 0:
        return INF
    product = LD(operations) * U64
    if not (product < 1):
        return INF
    return upward_from_long_double(product / (LD(1) - product))


def upward_add(left: float, right: float) -> float:
    return upward_from_long_double(LD(left) + LD(right))


def upward_multiply(left: float, right: float) -> float:
    if left < 0.0 or right < 0.0:
        return INF
    product = LD(left) * LD(right)
    if left > 0.0 and right > 0.0 and product == 0:
        return DENORM_MIN
    return upward_from_long_double(product)


def downward_add(left: float, right: float) -> float:
    if left < 0.0 or right < 0.0 or not math.isfinite(left) or not math.isfinite(right):
        return 0.0
    return downward_from_long_double(LD(left) + LD(right))


def downward_multiply(left: float, right: float) -> float:
    if left < 0.0 or right < 0.0 or not math.isfinite(left) or not math.isfinite(right):
        return 0.0
    return downward_from_long_double(LD(left) * LD(right))


def downward_divide(numerator: float, denominator: float) -> float:
    if not (numerator > 0.0) or not (denominator > 0.0) or not math.isfinite(numerator) \
            or not math.isfinite(denominator):
        return 0.0
    return downward_from_long_double(LD(numerator) / LD(denominator))


def downward_difference(left: float, right: float) -> float:
    if not (left > right) or not math.isfinite(left) or right < 0.0 or not math.isfinite(right):
        return 0.0
    return downward_from_long_double(LD(left) - LD(right))


def upward_divide(numerator: float, denominator: float) -> float:
    if numerator == 0.0:
        return 0.0
    if not (denominator > 0.0) or numerator < 0.0 or not math.isfinite(numerator) \
            or not math.isfinite(denominator):
        return INF
    quotient = LD(numerator) / LD(denominator)
    if quotient == 0:
        return DENORM_MIN
    return upward_from_long_double(quotient)


def _absolute_roundoff(operations: int) -> float:
    return upward_from_long_double(LD(operations) * LD(DENORM_MIN))


def positive_sum_upper_bound(stored_sum: float, operations: int) -> float:
    """floating_point_bounds.hpp:216-237."""
    if stored_sum < 0.0 or not math.isfinite(stored_sum) or operations < 0:
        return INF
    if stored_sum == 0.0:
        return 0.0
    gamma = binary64_gamma(operations)
    denominator = downward_difference(1.0, gamma)
    if not (denominator > 0.0):
        return INF
    return upward_divide(upward_add(stored_sum, _absolute_roundoff(operations)), denominator)


def positive_fma_square_sum_upper_bound(stored_sum: float, operations: int) -> float:
    """floating_point_bounds.hpp:243-259 (stored_sum == 0 does not imply 0)."""
    if stored_sum < 0.0 or not math.isfinite(stored_sum) or operations < 0:
        return INF
    gamma = binary64_gamma(operations)
    denominator = downward_difference(1.0, gamma)
    if not (denominator > 0.0):
        return INF
    return upward_divide(upward_add(stored_sum, _absolute_roundoff(operations)), denominator)


def positive_sum_lower_bound(stored_sum: float, operations: int) -> float:
    """floating_point_bounds.hpp:265-281."""
    if not (stored_sum > 0.0) or not math.isfinite(stored_sum) or operations < 0:
        return 0.0
    gamma = binary64_gamma(operations)
    denominator = upward_add(1.0, gamma)
    if not math.isfinite(denominator):
        return 0.0
    return downward_divide(downward_difference(stored_sum, _absolute_roundoff(operations)),
                           denominator)


def upward_scale_power_of_two(value: float, exponent: int) -> float:
    if value == 0.0:
        return 0.0
    if not (value > 0.0) or not math.isfinite(value):
        return INF
    try:
        scaled = math.ldexp(value, exponent)
    except OverflowError:
        return INF
    if math.isinf(scaled):
        return scaled
    if scaled == 0.0:
        return DENORM_MIN
    return float(np.nextafter(scaled, INF))


def downward_scale_power_of_two(value: float, exponent: int) -> float:
    if not (value > 0.0) or not math.isfinite(value):
        return 0.0
    try:
        scaled = math.ldexp(value, exponent)
    except OverflowError:
        return DBL_MAX
    if math.isinf(scaled):
        return DBL_MAX
    if scaled == 0.0:
        return 0.0
    return float(np.nextafter(scaled, 0.0))


def upward_sqrt(value: float) -> float:
    """floating_point_bounds.hpp:383-392."""
    if value < 0.0 or not math.isfinite(value):
        return value if value == INF else INF
    return upward_from_long_double(np.sqrt(LD(value)) * (LD(1) + LD(2) * LD_EPS))


def downward_sqrt(value: float) -> float:
    """floating_point_bounds.hpp:394-401."""
    if not (value > 0.0) or not math.isfinite(value):
        return 0.0
    return downward_from_long_double(np.sqrt(LD(value)) * (LD(1) - LD(2) * LD_EPS))


def norm_interval_from_parts(maximum: float, scaled_square_sum: float, count: int):
    """Interval from (max |v|, sum of squares of v scaled by 2**-e, length).

    floating_point_bounds.hpp:347-375, after the max/scale/sum passes.  The
    GPU norm kernels return the same three quantities, so tests can feed the
    device partials through this function.
    """
    if not math.isfinite(maximum) or not math.isfinite(scaled_square_sum):
        return (0.0, INF)
    if maximum == 0.0:
        return (0.0, 0.0)
    _, rescale_exponent = math.frexp(maximum)
    squared_lower = positive_sum_lower_bound(scaled_square_sum, count)
    squared_upper = positive_sum_upper_bound(scaled_square_sum, count)
    scaled_lower = downward_sqrt(squared_lower)
    scaled_upper = upward_sqrt(squared_upper)
    scaling_error = upward_from_long_double(np.sqrt(LD(count)) * LD(DENORM_MIN))
    return (downward_scale_power_of_two(downward_difference(scaled_lower, scaling_error),
                                        rescale_exponent),
            upward_scale_power_of_two(upward_add(scaled_upper, scaling_error), rescale_exponent))


def norm_interval(values) -> tuple:
    """Two-pass scaled norm interval (floating_point_bounds.hpp:325-376)."""
    v = np.asarray(values, dtype=np.float64).ravel()
    if v.size == 0:
        return (0.0, 0.0)
    a = np.abs(v)
    if not np.all(np.isfinite(a)):
        return (0.0, INF)
    maximum = float(a.max())
    if maximum == 0.0:
        return (0.0, 0.0)
    _, e = math.frexp(maximum)
    scaled = np.ldexp(a, -e)
    s = float(np.dot(scaled, scaled))
    return norm_interval_from_parts(maximum, s, int(v.size))


# Vectorised forms used by the operator constructor scans (same formulas
# applied elementwise; numpy longdouble arrays carry the x87 arithmetic).

def _upward_from_ld_array(values: np.ndarray) -> np.ndarray:
    values = values.astype(LD)
    out = np.full(values.shape, INF)
    ok = (values >= 0) & np.isfinite(values) & (values <= LD(DBL_MAX))
    zero = ok & (values == 0)
    nz = ok & ~zero
    r = values[nz].astype(np.float64)
    low = r.astype(LD) < values[nz]
    r = np.where(low, np.nextafter(r, INF), r)
    out[nz] = np.nextafter(r, INF)
    out[zero] = 0.0
    return out


def _downward_from_ld_array(values: np.ndarray) -> np.ndarray:
    values = values.astype(LD)
    out = np.zeros(values.shape)
    ok = (values > 0) & np.isfinite(values)
    big = ok & (values > LD(DBL_MAX))
    nz = ok & ~big
    r = values[nz].astype(np.float64)
    high = r.astype(LD) > values[nz]
    r = np.where(high, np.nextafter(r, 0.0), r)
    out[nz] = np.nextafter(r, 0.0)
    out[big] = DBL_MAX
    return out


def binary64_gamma_array(operations: np.ndarray) -> np.ndarray:
    ops = np.asarray(operations).astype(LD)
    product = ops * U64
    out = np.full(ops.shape, INF)
    ok = (ops >= 0) & (product < 1)
    out[ok] = _upward_from_ld_array(product[ok] / (LD(1) - product[ok]))
    return out


def positive_sum_upper_bound_array(stored: np.ndarray, ops: np.ndarray) -> np.ndarray:
    """Elementwise positive_sum_upper_bound (stored==0 -> 0)."""
    stored = np.asarray(stored, dtype=np.float64)
    ops = np.asarray(ops, dtype=np.int64)
    out = _positive_upper_common(stored, ops)
    out[(stored == 0.0) & (ops >= 0)] = 0.0
    return out


def positive_fma_square_sum_upper_bound_array(stored, ops) -> np.ndarray:
    return _positive_upper_common(np.asarray(stored, dtype=np.float64),
                                  np.asarray(ops, dtype=np.int64))


def _positive_upper_common(stored: np.ndarray, ops: np.ndarray) -> np.ndarray:
    out = np.full(stored.shape, INF)
    valid = (stored >= 0.0) & np.isfinite(stored) & (ops >= 0)
    gamma = binary64_gamma_array(ops)
    # downward_difference(1, gamma)
    denom = np.zeros(stored.shape)
    dd_ok = (gamma < 1.0) & np.isfinite(gamma)
    denom[dd_ok] = _downward_from_ld_array(LD(1) - gamma[dd_ok].astype(LD))
    abs_round = _upward_from_ld_array(ops.astype(LD) * LD(DENORM_MIN))
    num = _upward_from_ld_array(stored.astype(LD) + abs_round.astype(LD))
    good = valid & (denom > 0.0)
    q = num[good].astype(LD) / denom[good].astype(LD)
    res = _upward_from_ld_array(q)
    res[(num[good] != 0.0) & (q == 0)] = DENORM_MIN
    res[num[good] == 0.0] = 0.0
    # upward_divide returns inf for a nonfinite numerator
    res[~np.isfinite(num[good])] = INF
    out[good] = res
    return out


def positive_sum_lower_bound_array(stored, ops) -> np.ndarray:
    stored = np.asarray(stored, dtype=np.float64)
    ops = np.asarray(ops, dtype=np.int64)
    out = np.zeros(stored.shape)
    valid = (stored > 0.0) & np.isfinite(stored) & (ops >= 0)
    gamma = binary64_gamma_array(ops)
    denom = _upward_from_ld_array(LD(1) + gamma.astype(LD))
    abs_round = _upward_from_ld_array(ops.astype(LD) * LD(DENORM_MIN))
    diff = np.zeros(stored.shape)
    dok = valid & (stored > abs_round) & np.isfinite(abs_round)
    diff[dok] = _downward_from_ld_array(stored[dok].astype(LD) - abs_round[dok].astype(LD))
    good = valid & np.isfinite(denom) & (diff > 0.0) & (denom > 0.0)
    out[good] = _downward_from_ld_array(diff[good].astype(LD) / denom[good].astype(LD))
    return out


def downward_sqrt_array(values: np.ndarray) -> np.ndarray:
    values = np.asarray(values, dtype=np.float64)
    out = np.zeros(values.shape)
    ok = (values > 0.0) & np.isfinite(values)
    out[ok] = _downward_from_ld_array(np.sqrt(values[ok].astype(LD)) * (LD(1) - LD(2) * LD_EPS))
    return out
