"""The engine's host bound arithmetic (csrc/engine/fp_bounds.hpp, the
restatement of the reference's floating_point_bounds.hpp:12-403) against
exact rational arithmetic: every upward result is at or above the exact
value and at most two binary64 steps above it, every downward result at or
below and at most two steps below, gamma_k encloses k u / (1 - k u), and
the norm interval built from (max |v|, scaled square sum, n) encloses the
exact 2-norm of the vector the statistics came from."""
import math
import random
import shutil
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "native" / "fp_bounds_driver.cpp"


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("fp") / "fp_bounds_driver"
    subprocess.run([gxx, "-std=c++20", "-O2", "-o", str(exe), str(DRIVER)], check=True)
    return exe


def _run(exe, lines):
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return out.stdout.split("\n")[:len(lines)]


def _steps_above(value: float, exact: Fraction) -> int:
    """Number of binary64 steps from the largest double <= exact up to value
    (negative when value is below exact)."""
    if Fraction(value) >= exact:
        k, v = 0, value
        while Fraction(np.nextafter(v, -math.inf)) >= exact:
            v = float(np.nextafter(v, -math.inf))
            k += 1
            if k > 8:
                break
        return k
    return -1


def _steps_below(value: float, exact: Fraction) -> int:
    if Fraction(value) <= exact:
        k, v = 0, value
        while Fraction(np.nextafter(v, math.inf)) <= exact:
            v = float(np.nextafter(v, math.inf))
            k += 1
            if k > 8:
                break
        return k
    return -1


def _samples(rng, n):
    out = []
    for _ in range(n):
        e = rng.randint(-40, 40)
        out.append(math.ldexp(rng.random() + 0.5, e))
    return out + [1.0, 3.0, 0.1, 1e-300, 2.0 ** -1060, 1e300]


def test_binary_operations_round_outward(driver):
    rng = random.Random(7)
    xs = _samples(rng, 60)
    cases = []
    for i in range(0, len(xs) - 1):
        a, b = xs[i], xs[i + 1]
        cases += [("up_add", a, b, Fraction(a) + Fraction(b), "up"),
                  ("down_add", a, b, Fraction(a) + Fraction(b), "down"),
                  ("up_mul", a, b, Fraction(a) * Fraction(b), "up"),
                  ("down_mul", a, b, Fraction(a) * Fraction(b), "down"),
                  ("up_div", a, b, Fraction(a) / Fraction(b), "up"),
                  ("down_div", a, b, Fraction(a) / Fraction(b), "down")]
        hi, lo = max(a, b), min(a, b)
        if hi > lo:
            cases.append(("down_diff", hi, lo, Fraction(hi) - Fraction(lo), "down"))
    res = _run(driver, [f"{op} {float(a).hex()} {float(b).hex()}" for op, a, b, _, _ in cases])
    for (op, a, b, exact, side), line in zip(cases, res):
        v = float.fromhex(line)
        if exact > Fraction(np.finfo(np.float64).max) or (side == "up" and v == math.inf):
            continue
        if side == "up":
            k = _steps_above(v, exact)
            assert 0 <= k <= 2, (op, a, b, v, k)
        else:
            k = _steps_below(v, exact)
            assert 0 <= k <= 2 or v == 0.0 and exact >= 0, (op, a, b, v, k)


def test_sqrt_and_gamma(driver):
    rng = random.Random(3)
    xs = _samples(rng, 40)
    lines = [f"up_sqrt {x.hex()} 0x0p+0" for x in xs] + [f"down_sqrt {x.hex()} 0x0p+0" for x in xs]
    ks = [1, 2, 3, 7, 64, 1000, 2 ** 20, 10 ** 12]
    lines += [f"gamma64 {float(k).hex()} 0x0p+0" for k in ks]
    res = [float.fromhex(s) for s in _run(driver, lines)]
    n = len(xs)
    for x, up, dn in zip(xs, res[:n], res[n:2 * n]):
        assert Fraction(up) ** 2 >= Fraction(x) and Fraction(dn) ** 2 <= Fraction(x)
        assert up <= math.sqrt(x) * (1 + 1e-15) and dn >= math.sqrt(x) * (1 - 1e-15)
    u = Fraction(1, 2 ** 53)
    for k, g in zip(ks, res[2 * n:]):
        exact = k * u / (1 - k * u)
        assert 0 <= _steps_above(g, exact) <= 2, (k, g)


def test_norm_interval_encloses_exact_norm(driver):
    rng = np.random.default_rng(11)
    lines, exact = [], []
    for n, scale in [(1, 1.0), (5, 1e-3), (1000, 1.0), (4096, 1e200), (777, 1e-200), (100000, 3.0)]:
        v = rng.standard_normal(n) * scale
        v[rng.integers(0, n)] *= 16.0
        mx = float(np.max(np.abs(v)))
        _, e = math.frexp(mx)
        scaled = [math.ldexp(float(t), -e) for t in v]
        s = 0.0
        for t in scaled:
            s += t * t
        lines.append(f"norm {mx.hex()} {s.hex()} {n}")
        exact.append(sum(Fraction(float(t)) ** 2 for t in v))
    res = _run(driver, lines)
    for line, ex in zip(res, exact):
        lo, hi = (float.fromhex(w) for w in line.split())
        assert Fraction(lo) ** 2 <= ex <= Fraction(hi) ** 2, (lo, hi)
        assert hi <= lo * (1 + 1e-9)
