"""Counter-addressed Gaussian stream (oracle).

Follows src/algorithms/mathematics/random/gaussian.cpp:13-82:
key = splitmix64(seed) ^ splitmix64(stream + 0xd2b74407b1ce6e93);
u1, u2 = open_uniform(splitmix64(key ^ 2i)), open_uniform(splitmix64(key ^ (2i+1)));
z = sqrt(-2 ln u1) * cos(2 pi u2), with i = column * rows + row.

Pinned by tests/test_oracle_rng.py: splitmix64 against the published
SplitMix64 known-answer outputs (state 0 and state 1234567 of the public
reference generator), open_uniform against its closed form on edge bit
patterns, and the Gaussian stage against a scalar math.log/math.cos
evaluation plus moment checks.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
STREAM_SALT = 0xD2B74407B1CE6E93
MASK64 = (1 << 64) - 1


def splitmix64(value):
    """Vectorised splitmix64 finaliser (gaussian.cpp:13-18); wraps mod 2**64."""
    with np.errstate(over="ignore"):
        v = np.asarray(value, dtype=np.uint64) + GOLDEN
        v = (v ^ (v >> np.uint64(30))) * M1
        v = (v ^ (v >> np.uint64(27))) * M2
        return v ^ (v >> np.uint64(31))


def splitmix64_int(value: int) -> int:
    """Scalar Python-int splitmix64 used as an independent check."""
    v = (value + 0x9E3779B97F4A7C15) & MASK64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK64
    return v ^ (v >> 31)


def open_uniform(bits):
    """gaussian.cpp:20-23: ((bits >> 11) + 1) * 2**-53, in (0, 1]."""
    b = np.asarray(bits, dtype=np.uint64)
    return ((b >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)


def stream_key(seed: int, stream: int) -> np.uint64:
    s = splitmix64(np.uint64(seed & MASK64))
    t = splitmix64(np.uint64((stream + STREAM_SALT) & MASK64))
    return np.uint64(s ^ t)


def gaussian_at(seed: int, stream: int, index) -> np.ndarray:
    """gaussian.cpp:25-35, vectorised over ``index`` (uint64)."""
    key = stream_key(seed, stream)
    idx = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        two_i = idx * np.uint64(2)
    u1 = open_uniform(splitmix64(key ^ two_i))
    u2 = open_uniform(splitmix64(key ^ (two_i + np.uint64(1))))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def gaussian_matrix(rows: int, cols: int, seed: int, stream: int, first_column: int = 0):
    """gaussian.cpp:39-55: column-major rows x cols block, index = col*rows + row."""
    if rows == 0 or cols == 0:
        return np.zeros((rows, cols))
    col = np.arange(cols, dtype=np.uint64) + np.uint64(first_column)
    row = np.arange(rows, dtype=np.uint64)
    with np.errstate(over="ignore"):
        index = col[None, :] * np.uint64(rows) + row[:, None]
    return gaussian_at(seed, stream, index)
