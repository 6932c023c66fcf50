"""Counter-addressed synthetic CSR generator (oracle; test infrastructure).

CPU restatement of the engine's device generator for the batched benchmark
inputs.  The generator has no counterpart in the reference; its counters use
the splitmix64 / open-uniform / Box-Muller chain of
src/algorithms/mathematics/random/gaussian.cpp:13-35 (restated in
oracle/rng.py):

  stream 1, counter r             row length lo + floor(u * (hi - lo + 1)),
                                  capped at cols
  stream 2, counter r * 2^16 + a  column draw a of row r: floor(u * cols);
                                  repeats rejected, at most 2^16 draws, then
                                  the smallest unused column
  stream 3, counter e             entry e (rows in order, columns ascending):
                                  N(0,1) by Box-Muller
  stream 4, counter r             row scale 10^(log10 lo + u * (log10 hi - log10 lo))

Pinned by tests/test_oracle_csr_gen.py: the vectorised integer path against
a scalar Python-int evaluation of splitmix64 for the first rows, the CSR
invariants (row lengths in range, ascending distinct columns below cols),
a chi-square test of column uniformity, and the value stage against
rng.gaussian_at at the entry counters.
"""
from __future__ import annotations

import numpy as np

from . import rng


def _below(key: np.uint64, counters, bound: int) -> np.ndarray:
    u = rng.open_uniform(rng.splitmix64(np.uint64(key) ^ np.asarray(counters, dtype=np.uint64)))
    v = (u * float(bound)).astype(np.int64)
    return np.minimum(v, bound - 1)


def row_lengths(rows: int, cols: int, lo: int, hi: int, seed: int) -> np.ndarray:
    k = lo + _below(rng.stream_key(seed, 1), np.arange(rows, dtype=np.uint64), hi - lo + 1)
    return np.minimum(k, cols)


def row_columns(r: int, k: int, cols: int, seed: int) -> np.ndarray:
    key = rng.stream_key(seed, 2)
    chosen: list[int] = []
    seen: set[int] = set()
    a = 0
    while len(chosen) < k:
        if a < 65536:
            v = int(_below(key, [(r << 16) + a], cols)[0])
            a += 1
            if v in seen:
                continue
        else:
            v = 0
            while v in seen:
                v += 1
        chosen.append(v)
        seen.add(v)
    return np.array(sorted(chosen), dtype=np.int32)


def csr_generate(rows: int, cols: int, lo: int, hi: int, seed: int, scale_range=(1.0, 1.0)):
    """(row_ptr int64, col_idx int32, values float64) of the generated CSR."""
    k = row_lengths(rows, cols, lo, hi, seed)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(k, out=rp[1:])
    ci = np.empty(int(rp[-1]), dtype=np.int32)
    for r in range(rows):
        ci[rp[r]:rp[r + 1]] = row_columns(r, int(k[r]), cols, seed)
    vals = rng.gaussian_at(seed, 3, np.arange(int(rp[-1]), dtype=np.uint64))
    s_lo, s_hi = scale_range
    if not (s_lo == 1.0 and s_hi == 1.0):
        l0, l1 = np.log10(s_lo), np.log10(s_hi)
        u = rng.open_uniform(rng.splitmix64(rng.stream_key(seed, 4) ^ np.arange(rows, dtype=np.uint64)))
        scale = 10.0 ** (l0 + u * (l1 - l0))
        vals = vals * np.repeat(scale, k)
    return rp, ci, vals
