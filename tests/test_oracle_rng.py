"""Pins oracle/rng.py (reference random/gaussian.cpp:13-55)."""
import math

import numpy as np

from oracle import rng


def test_splitmix64_published_known_answers():
    # SplitMix64 reference outputs: state 0, and the sequence from state 1234567
    assert rng.splitmix64_int(0) == 0xE220A8397B1DCDAF
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                4593380528125082431, 16408922859458223821]
    state = 1234567
    for e in expected:
        assert rng.splitmix64_int(state) == e
        assert int(rng.splitmix64(np.uint64(state))) == e
        state = (state + 0x9E3779B97F4A7C15) % 2**64


def test_open_uniform_edges():
    # ((bits >> 11) + 1) * 2^-53: never 0, reaches exactly 1
    assert rng.open_uniform(np.uint64(0)) == 2.0 ** -53
    assert rng.open_uniform(np.uint64(2**64 - 1)) == 1.0
    assert rng.open_uniform(np.uint64(2**11)) == 2 * 2.0 ** -53


def test_gaussian_matches_scalar_box_muller_and_layout():
    seed, stream, rows, cols, first = 7, 3, 5, 4, 2
    g = rng.gaussian_matrix(rows, cols, seed, stream, first)
    key = rng.splitmix64_int(seed) ^ rng.splitmix64_int(stream + 0xD2B74407B1CE6E93)
    for c in range(cols):
        for r in range(rows):
            i = (first + c) * rows + r
            u1 = ((rng.splitmix64_int(key ^ (2 * i)) >> 11) + 1) * 2.0 ** -53
            u2 = ((rng.splitmix64_int(key ^ (2 * i + 1)) >> 11) + 1) * 2.0 ** -53
            z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
            assert g[r, c] == z


def test_gaussian_moments_and_counter_addressing():
    g = rng.gaussian_matrix(20000, 5, 11, 0, 0)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1.0) < 0.02
    # column j of a block started at first_col equals column first_col + j
    a = rng.gaussian_matrix(20000, 5, 11, 0, 0)
    b = rng.gaussian_matrix(20000, 2, 11, 0, 3)
    assert np.array_equal(a[:, 3:5], b)
