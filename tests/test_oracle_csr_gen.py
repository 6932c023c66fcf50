"""Pins the oracle's synthetic CSR generator (oracle/csr_gen.py)."""
import math

import numpy as np

from oracle import csr_gen as G
from oracle import rng


def _below_scalar(seed, stream, counter, bound):
    key = rng.splitmix64_int(seed & rng.MASK64) ^ rng.splitmix64_int((stream + rng.STREAM_SALT) & rng.MASK64)
    bits = rng.splitmix64_int(key ^ counter)
    u = ((bits >> 11) + 1) / 9007199254740992.0
    return min(int(u * bound), bound - 1)


def test_integer_path_matches_scalar_splitmix():
    rows, cols, lo, hi, seed = 40, 700, 4, 32, 12345
    k = G.row_lengths(rows, cols, lo, hi, seed)
    for r in range(rows):
        assert k[r] == lo + _below_scalar(seed, 1, r, hi - lo + 1)
        # first accepted columns follow the scalar draws with repeats rejected
        want, a = [], 0
        while len(want) < k[r]:
            v = _below_scalar(seed, 2, (r << 16) + a, cols)
            a += 1
            if v not in want:
                want.append(v)
        assert list(G.row_columns(r, int(k[r]), cols, seed)) == sorted(want)


def test_csr_invariants_and_value_stage():
    rows, cols, lo, hi, seed = 3000, 900, 4, 32, 7
    rp, ci, vals = G.csr_generate(rows, cols, lo, hi, seed)
    k = np.diff(rp)
    assert k.min() >= lo and k.max() <= hi and rp[0] == 0 and rp[-1] == ci.size == vals.size
    for r in range(rows):
        c = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < cols
    np.testing.assert_array_equal(vals, rng.gaussian_at(seed, 3, np.arange(rp[-1], dtype=np.uint64)))
    # lengths uniform over [lo, hi]; columns uniform over [0, cols)
    hist = np.bincount(k - lo, minlength=hi - lo + 1)
    exp = rows / (hi - lo + 1)
    assert np.sum((hist - exp) ** 2 / exp) < 70.0  # chi-square, 28 dof, p ~ 2e-5
    hc = np.bincount(ci, minlength=cols)
    exp_c = ci.size / cols
    chi = np.sum((hc - exp_c) ** 2 / exp_c)
    assert abs(chi - (cols - 1)) < 6.0 * math.sqrt(2.0 * (cols - 1))


def test_row_lengths_cap_and_scales():
    rp, ci, vals = G.csr_generate(50, 10, 8, 20, 3, scale_range=(1e-3, 1.0))
    k = np.diff(rp)
    assert k.max() <= 10 and k.min() >= 8
    for r in range(50):
        if k[r] == 10:
            assert list(ci[rp[r]:rp[r + 1]]) == list(range(10))
    base = rng.gaussian_at(3, 3, np.arange(rp[-1], dtype=np.uint64))
    ratio = vals / base
    per_row = [ratio[rp[r]:rp[r + 1]] for r in range(50)]
    for r, q in enumerate(per_row):
        assert np.allclose(q, q[0], rtol=1e-15) and 1e-3 * (1 - 1e-12) <= q[0] <= 1.0
