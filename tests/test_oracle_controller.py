"""Pins oracle/schedule.py and oracle/controller.py (reference
certificate_checkpoint_schedule.hpp, amfls_internal.hpp) by brute force and
hand-computed cases."""
import math

from oracle import controller as C
from oracle import schedule as S


def test_static_checkpoints_enumeration():
    pow2 = {1 << j for j in range(20)}
    mid = {3 << j for j in range(20)}
    for it in range(0, 5000):
        expect0 = it <= 1 or it in pow2 or it in mid
        expect1 = it <= 1 or it in pow2
        assert S.is_certificate_checkpoint(it, 0.0) == expect0
        assert S.is_certificate_checkpoint(it, 1e-3) == expect1


def test_predicted_crossing_log_linear():
    # 1e-2 at 8, 1e-4 at 16: two decades per 8 levels; reaching 3e-8 needs
    # log10(1e-4/3e-8) = 3.52 decades = 14.09 levels -> ceil -> 16 + 15 = 31
    assert S.predicted_crossing(1e-2, 1e-4, 8, 16, 3e-8) == 31
    assert S.predicted_crossing(1e-2, 1e-4, 8, 16, 1e-4) == S.INT_MAX  # already at tolerance
    assert S.predicted_crossing(1e-4, 1e-2, 8, 16, 1e-8) == S.INT_MAX  # no progress
    # a non-integral crossing rounds up
    assert S.predicted_crossing(1.0, 0.5, 0, 1, 0.3) == 1 + math.ceil(math.log(0.5 / 0.3) / math.log(2.0))


def test_dyadic_increment_brute_force():
    for active in range(0, 9):
        for gen in range(0, 9):
            for mx in range(0, 12):
                for avail in range(0, 12):
                    got = C.next_dyadic_active_width_increment(active, gen, mx, avail)
                    if active <= 0 or mx <= gen or active > gen + 1 or avail <= 0:
                        assert got == 0
                        continue
                    max_active = active + mx - gen
                    nxt = min(2 * active, max_active)
                    req = nxt - active
                    assert got == (req if 0 < req <= avail else 0)


def test_feedback_feasible_brute_force():
    for rank in range(0, 10):
        for limit in range(rank, 12):
            for frontier in range(0, 5):
                for levels in range(1, 4):
                    got = C.maximum_feedback_feasible_width_increment(rank, limit, frontier, levels)
                    best = 0
                    for inc in range(0, 20):
                        if (frontier + inc) * levels <= limit - rank:
                            best = inc
                    assert got == best


def test_cost_gate_hand_case():
    # width 1 -> 2, 8 matched columns: incumbent 8 levels, target 4 levels
    op = [1.0, 1.0]
    loc = [0.1, 0.1]
    d = C.make_horizon_width_candidate(1, 8, 0, 100, 0, 100, 1, op, loc, 2, 1.0, 2.0)
    assert (d.incumbent_levels, d.target_levels) == (8, 4)
    # incumbent 8*(1+0.1)=8.8; target 4*(1*2+0.4)+0.4 = 10.0 -> not admissible
    assert math.isclose(d.incumbent_estimated_seconds, 8.8)
    assert math.isclose(d.target_estimated_seconds, 10.0)
    assert not d.candidate
    # cheaper batched operator cost (1.0 vs 1.0): target 4*(1+0.4)+0.4 = 6.0 < 8.8
    d = C.make_horizon_width_candidate(1, 8, 0, 100, 0, 100, 1, op, loc, 2, 1.0, 1.0)
    assert math.isclose(d.target_estimated_seconds, 6.0) and d.candidate


def test_progress_history_shift_and_pool():
    h = C.ProgressHistory()
    for prev, cur in ((1.0, 0.1), (0.1, 0.01), (0.01, 0.005)):
        C.record_deepen_progress(h, prev, cur, 2, 2, 2, 1.0, 1.0)
    assert h.same_width_level_interval_count == 2
    assert math.isclose(h.prior_same_width_log_contraction, math.log(10.0))
    assert h.prior_same_width_levels == 2
    assert math.isclose(h.same_width_log_contractions[1], math.log(2.0))
    C.record_deepen_progress(h, 0.01, 0.001, 2, 4, 1)  # width changed -> reset
    assert h.same_width_level_interval_count == 0 and h.active_left_width == -1
