"""Adaptive-width controller helpers (oracle).

Follows src/algorithms/amfls/amfls_internal.hpp:14-674 (ProgressHistory,
record_deepen_progress, next_dyadic_active_width_increment,
maximum_feedback_feasible_width_increment, apply_width_cost_gate,
make_horizon_width_candidate, make_initial_width_candidate and the
MatchedHorizonPlan helpers).  long double arithmetic uses numpy.longdouble.

Pinned by tests/test_oracle_controller.py: brute-force enumeration of the
dyadic increment and feasibility helpers over small integer grids against
their defining properties, hand-computed cost-gate decisions, and the
history shift/pool invariants; and by tests/test_oracle_pins.py: hand-worked
tables for record_deepen_progress, make_initial_width_candidate (scalar
slowdown, costlier target, block stall, guards) and the MatchedHorizonPlan
life cycle.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

INF = math.inf
INT_MAX = 2**31 - 1
DBL_MAX = float(np.finfo(np.float64).max)
LD = np.longdouble


@dataclass
class ProgressHistory:
    same_width_log_contractions: list = field(default_factory=lambda: [0.0, 0.0])
    same_width_level_log_contractions: list = field(default_factory=lambda: [0.0, 0.0])
    same_width_interval_levels: list = field(default_factory=lambda: [0, 0])
    same_width_level_operator_seconds: list = field(default_factory=lambda: [INF, INF])
    same_width_level_local_seconds: list = field(default_factory=lambda: [INF, INF])
    same_width_level_interval_count: int = 0
    prior_same_width_log_contraction: float = 0.0
    prior_same_width_levels: int = 0
    active_left_width: int = -1


def reset_progress_history(h: ProgressHistory) -> None:
    h.__init__()


def record_deepen_progress(h: ProgressHistory, previous: float, current: float,
                           start_width: int = 0, end_width: int = 0, levels: int = 1,
                           op_seconds: float = INF, local_seconds: float = INF) -> None:
    """amfls_internal.hpp:35-144."""
    if start_width <= 0 or end_width <= 0 or start_width != end_width:
        reset_progress_history(h)
        return
    if h.active_left_width != start_width:
        reset_progress_history(h)
        h.active_left_width = start_width
    if levels <= 0 or not (previous > 0.0) or not (current > 0.0) or \
            not math.isfinite(previous) or not math.isfinite(current):
        return
    progress = (np.log(LD(previous)) - np.log(LD(current))) if previous > current else LD(0)
    if not (progress >= 0) or not np.isfinite(progress):
        return
    level_lc = progress / LD(levels)
    level_lc_d = float(min(level_lc, LD(DBL_MAX))) if (level_lc > 0 and np.isfinite(level_lc)) else 0.0
    finite_cost = (op_seconds >= 0.0 and local_seconds >= 0.0 and math.isfinite(op_seconds)
                   and math.isfinite(local_seconds) and op_seconds + local_seconds > 0.0
                   and math.isfinite(op_seconds + local_seconds))
    lvl_op = op_seconds / levels if finite_cost else INF
    lvl_loc = local_seconds / levels if finite_cost else INF
    if h.same_width_level_interval_count < 2:
        i = h.same_width_level_interval_count
        h.same_width_log_contractions[i] = float(progress)
        h.same_width_level_log_contractions[i] = level_lc_d
        h.same_width_interval_levels[i] = levels
        h.same_width_level_operator_seconds[i] = lvl_op
        h.same_width_level_local_seconds[i] = lvl_loc
        h.same_width_level_interval_count += 1
    else:
        pooled = LD(h.prior_same_width_log_contraction) + LD(h.same_width_log_contractions[0])
        if h.prior_same_width_levels > INT_MAX - h.same_width_interval_levels[0] or \
                not np.isfinite(pooled) or pooled > LD(DBL_MAX):
            h.prior_same_width_log_contraction = INF
            h.prior_same_width_levels = INT_MAX
        else:
            h.prior_same_width_log_contraction = float(pooled)
            h.prior_same_width_levels += h.same_width_interval_levels[0]
        h.same_width_log_contractions = [h.same_width_log_contractions[1], float(progress)]
        h.same_width_level_log_contractions = [h.same_width_level_log_contractions[1], level_lc_d]
        h.same_width_interval_levels = [h.same_width_interval_levels[1], levels]
        h.same_width_level_operator_seconds = [h.same_width_level_operator_seconds[1], lvl_op]
        h.same_width_level_local_seconds = [h.same_width_level_local_seconds[1], lvl_loc]


def next_dyadic_active_width_increment(active: int, generated: int, max_generated: int,
                                       available: int) -> int:
    """amfls_internal.hpp:151-181."""
    if active <= 0 or generated < 0 or max_generated <= generated or \
            active > generated + 1 or available <= 0:
        return 0
    max_active = active + max_generated - generated
    nxt = max_active if active > max_active // 2 else min(2 * active, max_active)
    if nxt <= active:
        return 0
    req = nxt - active
    return req if (req <= available and req <= INT_MAX) else 0


def ceiling_ratio(num: int, den: int) -> int:
    if num <= 0 or den <= 0:
        return 0
    return 1 + (num - 1) // den


def maximum_feedback_feasible_width_increment(rank: int, limit: int, frontier: int,
                                              levels: int) -> int:
    """amfls_internal.hpp:193-207."""
    if rank < 0 or limit < rank or frontier < 0 or levels <= 0:
        return 0
    return max(0, (limit - rank) // levels - frontier)


@dataclass
class WidthDecision:
    recent_level_log_contraction: float = 0.0
    prior_level_log_contraction: float = 0.0
    remaining_log_contraction: float = 0.0
    forecast_levels: int = 0
    incumbent_levels: int = 0
    target_levels: int = 0
    target_remaining_levels: int = 0
    matched_horizon: int = 0
    incumbent_estimated_seconds: float = INF
    target_estimated_seconds: float = INF
    history_mature: bool = False
    prior_history_mature: bool = False
    recent_slowdown: bool = False
    persistent_slowdown: bool = False
    progress_stalled: bool = False
    long_horizon_pressure: bool = False
    widening_necessary: bool = False
    observed_costs_mature: bool = False
    width_depth_admissible: bool = True
    pass_reduction_admissible: bool = False
    stage_feasible: bool = False
    finite_current_certificate: bool = False
    finite_resource_boundary: bool = False
    candidate: bool = False


def apply_width_cost_gate(d: WidthDecision, active: int, target: int, op_s, loc_s,
                          samples: int, inc_cost: float, tgt_cost: float) -> None:
    """amfls_internal.hpp:238-325."""
    if d.incumbent_levels <= 0 or d.target_levels <= 0 or active <= 0 or target <= active or \
            samples != 2 or not (inc_cost > 0.0) or not (tgt_cost > 0.0) or \
            not math.isfinite(inc_cost) or not math.isfinite(tgt_cost):
        return
    ratio = LD(target) / LD(active)
    ratio2 = ratio * ratio
    op_ratio = LD(tgt_cost) / LD(inc_cost)
    inc_op_cost = LD(d.incumbent_levels) * LD(inc_cost)
    tgt_op_cost = LD(d.target_levels) * LD(tgt_cost)
    best_inc = LD(INF)
    worst_tgt = LD(0)
    finite = True
    for i in range(samples):
        o = LD(op_s[i])
        lo = LD(loc_s[i])
        if not (o >= 0) or not (lo >= 0) or not np.isfinite(o) or not np.isfinite(lo) or \
                not (o + lo > 0):
            finite = False
            break
        inc_sec = LD(d.incumbent_levels) * (o + lo)
        tgt_loc = lo * ratio2
        tgt_sec = LD(d.target_levels) * (o * op_ratio + tgt_loc) + tgt_loc
        if not np.isfinite(inc_sec) or not np.isfinite(tgt_sec):
            finite = False
            break
        best_inc = min(best_inc, inc_sec)
        worst_tgt = max(worst_tgt, tgt_sec)
    if finite:
        finite = best_inc <= LD(DBL_MAX) and worst_tgt <= LD(DBL_MAX)
        if finite:
            d.incumbent_estimated_seconds = float(best_inc)
            d.target_estimated_seconds = float(worst_tgt)
    d.observed_costs_mature = finite
    d.pass_reduction_admissible = (d.width_depth_admissible and d.target_levels < d.incumbent_levels
                                   and tgt_op_cost < inc_op_cost and d.observed_costs_mature
                                   and d.target_estimated_seconds < d.incumbent_estimated_seconds)


def make_horizon_width_candidate(inc: int, remaining_horizon: int, depth: int, max_depth: int,
                                 rank: int, limit: int, active: int, op_s, loc_s, samples: int,
                                 inc_cost: float = 1.0, tgt_cost: float = 1.0,
                                 require_balance: bool = False) -> WidthDecision:
    """amfls_internal.hpp:332-392."""
    d = WidthDecision()
    d.stage_feasible = inc > 0 and active > 0 and inc <= INT_MAX - active
    d.finite_resource_boundary = (depth >= 0 and max_depth >= depth and rank >= 0 and limit >= rank
                                  and active > 0 and remaining_horizon > 0
                                  and remaining_horizon <= limit - rank)
    if not d.stage_feasible or not d.finite_resource_boundary:
        return d
    target = active + inc
    d.target_remaining_levels = min(max_depth - depth, (limit - rank) // target)
    cap = d.target_remaining_levels * target
    d.matched_horizon = int(min(remaining_horizon, cap))
    d.incumbent_levels = ceiling_ratio(d.matched_horizon, active)
    d.target_levels = ceiling_ratio(d.matched_horizon, target)
    if d.target_levels < 2:
        return d
    d.width_depth_admissible = (not require_balance) or d.target_levels > target
    apply_width_cost_gate(d, active, target, op_s, loc_s, samples, inc_cost, tgt_cost)
    d.candidate = d.pass_reduction_admissible
    return d


def make_initial_width_candidate(inc: int, cert: float, tol: float, depth: int, max_depth: int,
                                 rank: int, limit: int, active: int, h: ProgressHistory,
                                 inc_cost: float = 1.0, tgt_cost: float = 1.0) -> WidthDecision:
    """amfls_internal.hpp:399-552."""
    d = WidthDecision()
    d.finite_current_certificate = (cert > 0.0 and tol > 0.0 and cert > tol and math.isfinite(cert)
                                    and math.isfinite(tol))
    d.history_mature = h.active_left_width == active and h.same_width_level_interval_count >= 2
    if not d.finite_current_certificate or not d.history_mature or depth < 0 or \
            max_depth < depth or rank < 0 or limit < rank or active <= 0 or inc <= 0 or \
            inc > INT_MAX - active:
        return d
    recent_levels = h.same_width_interval_levels[0] + h.same_width_interval_levels[1]
    recent_progress = LD(h.same_width_log_contractions[0]) + LD(h.same_width_log_contractions[1])
    if recent_levels <= 0 or not (recent_progress >= 0) or not np.isfinite(recent_progress):
        return d
    d.progress_stalled = bool(recent_progress == 0)
    d.recent_level_log_contraction = 0.0 if d.progress_stalled else \
        float(recent_progress / LD(recent_levels))
    d.remaining_log_contraction = math.log(cert) - math.log(tol)
    if (not d.progress_stalled and (not (d.recent_level_log_contraction > 0.0)
                                    or not math.isfinite(d.recent_level_log_contraction))) or \
            not (d.remaining_log_contraction > 0.0) or not math.isfinite(d.remaining_log_contraction):
        return d
    if not d.progress_stalled:
        raw = LD(d.remaining_log_contraction) / LD(d.recent_level_log_contraction)
        if not (raw > 0) or not np.isfinite(raw):
            return d
        d.forecast_levels = INT_MAX if raw >= LD(INT_MAX) else max(1, int(np.ceil(raw)))
    d.prior_history_mature = (h.prior_same_width_levels > 0 and h.prior_same_width_log_contraction > 0.0
                              and math.isfinite(h.prior_same_width_log_contraction))
    if d.prior_history_mature:
        d.prior_level_log_contraction = h.prior_same_width_log_contraction / h.prior_same_width_levels
        lc = h.same_width_level_log_contractions
        d.persistent_slowdown = (math.isfinite(d.prior_level_log_contraction) and lc[0] > 0.0
                                 and lc[1] > 0.0 and lc[0] < d.prior_level_log_contraction
                                 and lc[1] < d.prior_level_log_contraction)
    remaining_basis = limit - rank
    target = active + inc
    target_remaining_levels = min(max_depth - depth, remaining_basis // target)
    raw_horizon = target * target_remaining_levels if d.progress_stalled else active * d.forecast_levels
    matched = int(min(raw_horizon, remaining_basis))
    c = make_horizon_width_candidate(inc, matched, depth, max_depth, rank, limit, active,
                                     h.same_width_level_operator_seconds,
                                     h.same_width_level_local_seconds, 2, inc_cost, tgt_cost)
    c.finite_current_certificate = d.finite_current_certificate
    c.history_mature = d.history_mature
    c.prior_history_mature = d.prior_history_mature
    c.persistent_slowdown = d.persistent_slowdown
    c.progress_stalled = d.progress_stalled
    c.recent_level_log_contraction = d.recent_level_log_contraction
    c.prior_level_log_contraction = d.prior_level_log_contraction
    c.remaining_log_contraction = d.remaining_log_contraction
    c.forecast_levels = d.forecast_levels
    c.long_horizon_pressure = c.target_remaining_levels > 0 and (
        c.progress_stalled or d.forecast_levels > c.target_remaining_levels)
    lc = h.same_width_level_log_contractions
    positive_recent = lc[0] > 0.0 and lc[1] > 0.0
    c.recent_slowdown = positive_recent and lc[1] < 0.5 * lc[0]
    established_stall = c.progress_stalled and (active > 1 or c.prior_history_mature)
    scalar_needed = active == 1 and positive_recent and (
        c.recent_slowdown or (c.prior_history_mature and (c.persistent_slowdown
                                                         or c.long_horizon_pressure)))
    block_needed = active > 1 and (positive_recent or established_stall) and (
        c.persistent_slowdown or c.long_horizon_pressure or c.pass_reduction_admissible)
    c.widening_necessary = scalar_needed or block_needed
    c.candidate = c.candidate and c.widening_necessary
    return c


@dataclass
class MatchedHorizonPlan:
    end_basis_rank: int = -1
    origin_active_width: int = -1
    active_left_width: int = -1
    starting_certificate: float = INF
    level_operator_seconds: list = field(default_factory=lambda: [INF, INF])
    level_local_seconds: list = field(default_factory=lambda: [INF, INF])
    cost_sample_count: int = 0


def reset_matched_horizon_plan(p: MatchedHorizonPlan) -> None:
    p.__init__()


def remaining_matched_horizon(p: MatchedHorizonPlan, rank: int) -> int:
    if p.end_basis_rank < 0 or rank < 0 or rank >= p.end_basis_rank:
        return 0
    return p.end_basis_rank - rank


def extend_matched_horizon_to_basis_limit(p: MatchedHorizonPlan, rank: int, limit: int) -> bool:
    if p.end_basis_rank < 0 or rank < 0 or limit <= rank or rank < p.end_basis_rank:
        return False
    p.end_basis_rank = limit
    return True


def matched_horizon_progress_stalled(p: MatchedHorizonPlan, cert: float) -> bool:
    return (p.starting_certificate > 0.0 and cert > 0.0 and math.isfinite(p.starting_certificate)
            and math.isfinite(cert)
            and math.log(p.starting_certificate) - math.log(cert) < math.log(2.0))


def start_matched_horizon_plan(p: MatchedHorizonPlan, rank: int, limit: int, horizon: int,
                               origin: int, target: int, cert: float) -> bool:
    if rank < 0 or limit < rank or horizon <= 0 or horizon > limit - rank or origin <= 0 or \
            target <= 0 or not (cert > 0.0) or not math.isfinite(cert):
        reset_matched_horizon_plan(p)
        return False
    reset_matched_horizon_plan(p)
    p.end_basis_rank = rank + horizon
    p.origin_active_width = origin
    p.active_left_width = target
    p.starting_certificate = cert
    return True


def reset_matched_horizon_costs(p: MatchedHorizonPlan, active: int) -> None:
    p.active_left_width = active
    p.level_operator_seconds = [INF, INF]
    p.level_local_seconds = [INF, INF]
    p.cost_sample_count = 0


def record_matched_horizon_cost(p: MatchedHorizonPlan, start_width: int, end_width: int,
                                levels: int, op_seconds: float, local_seconds: float) -> bool:
    if p.end_basis_rank < 0 or start_width <= 0 or start_width != end_width or \
            start_width != p.active_left_width or levels <= 0 or op_seconds < 0.0 or \
            local_seconds < 0.0 or not math.isfinite(op_seconds) or \
            not math.isfinite(local_seconds) or not (op_seconds + local_seconds > 0.0) or \
            p.cost_sample_count >= 2:
        return False
    i = p.cost_sample_count
    p.level_operator_seconds[i] = op_seconds / levels
    p.level_local_seconds[i] = local_seconds / levels
    p.cost_sample_count += 1
    return True
