"""Pins oracle/fpbounds.py (reference floating_point_bounds.hpp) with exact
rational arithmetic: every lower bound is <= the exact value <= every upper."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import fpbounds as fp

U = Fraction(1, 2**53)


def test_gamma_closed_form():
    for k in (0, 1, 2, 10, 1000, 10**9):
        exact = Fraction(k) * U / (1 - Fraction(k) * U)
        g = fp.binary64_gamma(k)
        assert Fraction(g) >= exact
        assert g <= float(exact) * (1 + 4e-16) + 1e-300
    assert fp.binary64_gamma(-1) == math.inf
    assert fp.binary64_gamma(2**53) == math.inf


@pytest.mark.parametrize("seed", range(5))
def test_directed_arithmetic_encloses_exact(seed):
    r = np.random.default_rng(seed)
    for _ in range(300):
        a, b = (float(v) for v in np.abs(r.standard_normal(2)) * 10.0 ** r.integers(-30, 30, 2))
        A, B = Fraction(a), Fraction(b)
        assert Fraction(fp.upward_add(a, b)) >= A + B >= Fraction(fp.downward_add(a, b))
        assert Fraction(fp.upward_multiply(a, b)) >= A * B >= Fraction(fp.downward_multiply(a, b))
        assert Fraction(fp.upward_divide(a, b)) >= A / B >= Fraction(fp.downward_divide(a, b))
        if a > b:
            assert A - B >= Fraction(fp.downward_difference(a, b))
        up, lo = fp.upward_sqrt(a), fp.downward_sqrt(a)
        assert Fraction(up) ** 2 >= A >= Fraction(lo) ** 2


def test_special_cases():
    assert fp.upward_add(0.0, 0.0) == 0.0
    assert fp.upward_multiply(5e-324, 5e-324) >= 5e-324  # positive product never rounds to 0
    assert fp.downward_difference(1.0, 2.0) == 0.0
    assert fp.upward_divide(1.0, 0.0) == math.inf
    assert fp.upward_sqrt(math.inf) == math.inf
    assert fp.positive_sum_upper_bound(0.0, 10) == 0.0
    assert fp.positive_fma_square_sum_upper_bound(0.0, 10) > 0.0  # underflowed squares
    assert fp.norm_interval(np.array([0.0, 0.0])) == (0.0, 0.0)
    assert fp.norm_interval(np.array([1.0, np.nan]))[1] == math.inf


@pytest.mark.parametrize("scale", [1e-300, 1e-160, 1.0, 1e150, 1e300])
def test_norm_interval_encloses_exact_norm(scale):
    r = np.random.default_rng(1)
    v = r.standard_normal(257) * scale
    lo, hi = fp.norm_interval(v)
    exact_sq = sum(Fraction(float(x)) ** 2 for x in v)
    assert Fraction(lo) ** 2 <= exact_sq <= Fraction(hi) ** 2
    assert hi / lo < 1 + 1e-12 if lo > 0 else True


def test_subnormal_vector_interval():
    v = np.array([5e-324, 1e-310, -3e-320])
    lo, hi = fp.norm_interval(v)
    exact_sq = sum(Fraction(float(x)) ** 2 for x in v)
    assert Fraction(lo) ** 2 <= exact_sq <= Fraction(hi) ** 2


def test_sum_bounds_enclose_reordered_sums():
    r = np.random.default_rng(2)
    x = np.abs(r.standard_normal(1000))
    s = float(np.sum(x))
    exact = sum(Fraction(float(t)) for t in x)
    assert Fraction(fp.positive_sum_lower_bound(s, x.size)) <= exact <= \
        Fraction(fp.positive_sum_upper_bound(s, x.size))
    # vectorised forms agree with the scalar ones
    sums = np.abs(r.standard_normal(50)) * 100
    ops = r.integers(1, 500, 50)
    up = fp.positive_sum_upper_bound_array(sums, ops)
    lo = fp.positive_sum_lower_bound_array(sums, ops)
    for i in range(50):
        assert up[i] == fp.positive_sum_upper_bound(float(sums[i]), int(ops[i]))
        assert lo[i] == fp.positive_sum_lower_bound(float(sums[i]), int(ops[i]))
