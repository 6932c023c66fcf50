"""AMFLS driver, fixed-width RBGK and LSQR (oracle).

* ``solve_amfls``      -- src/algorithms/amfls/amfls.cpp:25-661
* ``solve_fixed_rbgk`` -- src/algorithms/fixed_rbgk/fixed_rbgk.cpp:18-209
* ``solve_lsqr``       -- src/algorithms/lsqr/lsqr.cpp:84-279

``clock`` may be a SyntheticClock so the timing-driven width decisions are
deterministic (the reference hard-wires std::chrono).

Pinned by tests/test_oracle_solvers.py: the README diagonal example
(x = (1,2,3,4)), agreement with numpy.linalg.lstsq / closed-form ridge
solutions on seeded problems, LSQR iterates against scipy.sparse.linalg.lsqr,
counter identities (search A* columns = sum of left input widths), and the
terminal statuses forced from the host (maximum_depth, maximum_basis_size,
NaN operator).
"""
from __future__ import annotations

import math

import numpy as np

from . import controller as C
from . import results as R
from .clock import WallClock
from .rng import gaussian_matrix
from .schedule import CertificateCheckpointSchedule
from .validation import Counted, failed_numerically, validate_common_inputs, \
    validate_original_candidate
from .workspace import Workspace, step_has_nonfinite_breakdown, \
    step_has_numerical_rank_truncation

INF = math.inf


def default_epochs(dimension: int) -> int:
    epochs, covered = 1, 1
    while covered < dimension and epochs < 30:
        covered *= 2
        epochs += 1
    return epochs


def next_certificate_checkpoint(depth: int, max_depth: int, schedule) -> int:
    cp = depth + 1
    while cp < max_depth and not schedule.should_evaluate(cp):
        cp += 1
    return cp


def active_certificate(res, lam) -> float:
    return res.relative_energy_error_upper_bound if lam > 0.0 else res.backward_error_upper_bound


def solve_amfls(op, b, regularization=0.0, tolerance=1e-8, failure_probability=1e-6,
                maximum_epochs=0, maximum_basis_size=0, seed=0, stream=0, maximum_depth=0,
                maximum_auxiliary_width=0, clock=None) -> R.LeastSquaresResult:
    """amfls.cpp:57-661."""
    clock = clock or WallClock()
    b = np.asarray(b, dtype=np.float64)
    validate_common_inputs(op, b, regularization, tolerance)
    if not math.isfinite(failure_probability) or not (0.0 < failure_probability < 1.0):
        raise ValueError("failure_probability must lie strictly between zero and one")
    if maximum_epochs < 0 or maximum_depth < 0 or maximum_auxiliary_width < 0 or \
            maximum_basis_size < 0:
        raise ValueError("work limits must be nonnegative")
    lam, tol = regularization, tolerance
    total_start = clock.now()
    stats = R.RunStatistics()
    ws = Workspace(op, b, lam, maximum_basis_size, stats, clock)
    dimension = min(op.rows, op.cols)
    max_depth = min(maximum_depth, dimension) if maximum_depth > 0 else dimension
    cap_aux = max(0, ws.basis_limit - 1)
    max_aux = min(maximum_auxiliary_width if maximum_auxiliary_width > 0 else cap_aux, cap_aux)
    max_epochs = maximum_epochs if maximum_epochs > 0 else default_epochs(dimension)

    st = dict(trace=[], evaluations=0, last_depth=-1, aux=0, result=None, best=None,
              best_cert=INF, has=False)
    sched = [CertificateCheckpointSchedule(lam, tol)]

    def record(cand: R.LeastSquaresResult, train: bool):
        if len(cand.trace) != 1:
            raise RuntimeError("AMFLS candidate must contain exactly one trace record")
        if cand.depth == st["last_depth"]:
            raise RuntimeError("AMFLS evaluated one checkpoint more than once")
        st["last_depth"] = cand.depth
        if train:
            sched[0].record_evaluation(cand.depth, cand)
        st["trace"].append(cand.trace[0])
        cand.trace = []
        st["evaluations"] += 1
        cert = active_certificate(cand, lam)
        better = (not st["has"]) or (math.isfinite(cert) and not math.isfinite(st["best_cert"])) \
            or cert < st["best_cert"]
        if better:
            st["best_cert"] = cert
            st["best"] = cand.copy()
        st["has"] = True
        st["result"] = cand

    def evaluate_checkpoint(train: bool):
        record(ws.evaluate(tol, st["evaluations"], st["aux"]), train)

    def finish(final: R.LeastSquaresResult):
        final = final.copy()
        final.iterations = ws.depth
        final.trace = list(st["trace"])
        stats.total_seconds = clock.now() - total_start
        final.statistics = R.RunStatistics(**vars(stats))
        return final

    def finish_at_limit(status, reason):
        final = (st["result"] if lam > 0.0 else st["best"]).copy()
        final.status, final.stop_reason = status, reason
        return finish(final)

    ws.initialize(None)
    if ws.can_advance() and max_depth > 0:
        ws.advance_level()
    evaluate_checkpoint(True)
    res = st["result"]
    if failed_numerically(res) or res.status == R.SUCCESS:
        return finish(res)
    truncated = step_has_numerical_rank_truncation(ws.audit)
    broke = step_has_nonfinite_breakdown(ws.audit)
    hist = C.ProgressHistory()
    plan = C.MatchedHorizonPlan()
    epoch = 0
    random_generated = 0
    bhat = {"init": False, "vec": None}

    def gaussian_orthogonal_to_rhs(cols, first):
        rnd = gaussian_matrix(op.rows, cols, seed, stream, first)
        if not bhat["init"]:
            nb = float(np.linalg.norm(b))
            if nb != 0.0:
                bhat["vec"] = b / nb
            bhat["init"] = True
        if bhat["vec"] is None:
            return rnd
        v = bhat["vec"]
        for c in range(cols):
            coef = float(np.dot(v, rnd[:, c]))
            rnd[:, c] -= coef * v
        return rnd

    feedback_pending = False
    fb_target = 0
    stage_start_rank = 0
    stage_origin = 0
    stage_target = 0
    while True:
        res = st["result"]
        if broke:
            return finish_at_limit(R.NUMERICAL_BREAKDOWN, R.STOP_NUMERICAL_BREAKDOWN)
        if ws.basis_rank >= ws.basis_limit:
            return finish_at_limit(R.BASIS_LIMIT, R.MAXIMUM_BASIS)
        if ws.depth >= max_depth:
            return finish_at_limit(R.WORK_LIMIT, R.MAXIMUM_DEPTH)
        if truncated:
            return finish_at_limit(R.PRECISION_LIMIT, R.STOP_PRECISION_LIMIT)
        interval_cert = active_certificate(res, lam)
        frontier = ws.active_width
        rem_basis = max(0, ws.basis_limit - ws.basis_rank)
        avail_inc = max(0, rem_basis - frontier)
        rem_aux = max(0, max_aux - st["aux"])
        width_cap = min(avail_inc, rem_aux)
        fb_available = max_depth - ws.depth >= 2
        fb_depth = ws.depth + 2 if fb_available else ws.depth
        levels_to_fb = fb_depth - ws.depth
        fb_feasible = C.maximum_feedback_feasible_width_increment(
            ws.basis_rank, ws.basis_limit, frontier, levels_to_fb) if fb_available else 0
        feasible_cap = min(width_cap, fb_feasible)
        if plan.end_basis_rank >= 0 and C.remaining_matched_horizon(plan, ws.basis_rank) == 0:
            C.reset_matched_horizon_plan(plan)
        rem_horizon = C.remaining_matched_horizon(plan, ws.basis_rank)
        horizon_active = rem_horizon > 0
        collecting = horizon_active and not feedback_pending and plan.cost_sample_count < 2
        ready = not feedback_pending and not collecting
        stage_available = ready and epoch < max_epochs
        next_inc = C.next_dyadic_active_width_increment(
            frontier, st["aux"], max_aux, feasible_cap) if stage_available else 0
        requested = 0
        requested_horizon = 0
        if stage_available and next_inc > 0:
            target_active = frontier + next_inc
            if horizon_active:
                d = C.make_horizon_width_candidate(
                    next_inc, rem_horizon, ws.depth, max_depth, ws.basis_rank, ws.basis_limit,
                    frontier, plan.level_operator_seconds, plan.level_local_seconds,
                    plan.cost_sample_count, op.relative_block_product_cost(frontier),
                    op.relative_block_product_cost(target_active), plan.origin_active_width > 1)
            else:
                d = C.make_initial_width_candidate(
                    next_inc, interval_cert, tol, ws.depth, max_depth, ws.basis_rank,
                    ws.basis_limit, frontier, hist, op.relative_block_product_cost(frontier),
                    op.relative_block_product_cost(target_active))
            if d.candidate:
                requested = next_inc
                if not horizon_active:
                    requested_horizon = d.matched_horizon
        if horizon_active and ready and requested == 0:
            C.reset_matched_horizon_plan(plan)

        if requested > 0:
            fb_target = fb_depth
            stage_start_rank = ws.basis_rank
            stage_origin = plan.origin_active_width if horizon_active else frontier
            stage_target = frontier + requested
            C.reset_progress_history(hist)
            sched[0] = CertificateCheckpointSchedule(lam, tol)
            rnd = gaussian_orthogonal_to_rhs(requested, random_generated)
            stats.gaussian_random_block_requests += 1
            stats.gaussian_random_columns += requested
            stats.gaussian_random_values += op.rows * requested
            random_generated += requested
            st["aux"] += requested
            epoch += 1
            ws.append_left_seed(rnd)
            advanced = ws.advance_level() if ws.can_advance() else False
            if advanced:
                if requested_horizon > 0:
                    C.start_matched_horizon_plan(plan, stage_start_rank, ws.basis_limit,
                                                 requested_horizon, stage_origin, stage_target,
                                                 interval_cert)
                else:
                    C.reset_matched_horizon_costs(plan, stage_target)
                truncated = step_has_numerical_rank_truncation(ws.audit)
                broke = step_has_nonfinite_breakdown(ws.audit)
                exhausted = C.remaining_matched_horizon(plan, ws.basis_rank) == 0
                terminal = ws.basis_rank >= ws.basis_limit or ws.depth >= max_depth or \
                    not ws.can_advance()
                if terminal or exhausted:
                    evaluate_checkpoint(False)
                    res = st["result"]
                    if failed_numerically(res) or res.status == R.SUCCESS:
                        return finish(res)
                reliable = (not truncated and not broke and ws.basis_rank > stage_start_rank
                            and ws.active_width == stage_target)
                extended = exhausted and C.matched_horizon_progress_stalled(
                    plan, active_certificate(st["result"], lam)) and \
                    C.extend_matched_horizon_to_basis_limit(plan, ws.basis_rank, ws.basis_limit)
                if not reliable or (exhausted and not extended):
                    C.reset_matched_horizon_plan(plan)
                    feedback_pending = False
                else:
                    feedback_pending = True
            else:
                truncated = step_has_numerical_rank_truncation(ws.audit)
                broke = step_has_nonfinite_breakdown(ws.audit)
                feedback_pending = False
                C.reset_matched_horizon_plan(plan)
            continue

        if ws.can_advance():
            start_depth = ws.depth
            start_rank = ws.basis_rank
            matched_cost = (not feedback_pending and
                            C.remaining_matched_horizon(plan, start_rank) > 0 and
                            plan.cost_sample_count < 2)
            if feedback_pending:
                target_depth = fb_target
            elif matched_cost:
                target_depth = min(start_depth + 1, max_depth)
            else:
                target_depth = next_certificate_checkpoint(start_depth, max_depth, sched[0])
            t0 = clock.now()
            a0, at0 = stats.a_seconds, stats.at_seconds
            while ws.can_advance() and ws.depth < target_depth:
                advanced = ws.advance_level()
                truncated = step_has_numerical_rank_truncation(ws.audit)
                broke = step_has_nonfinite_breakdown(ws.audit)
                if not advanced or truncated or broke:
                    break
            search_s = clock.now() - t0
            op_s = max(0.0, stats.a_seconds - a0 + stats.at_seconds - at0)
            loc_s = max(0.0, search_s - op_s)
            if ws.depth > start_depth:
                completes_pair = matched_cost and plan.cost_sample_count == 1
                exhausts = (feedback_pending or matched_cost) and \
                    C.remaining_matched_horizon(plan, ws.basis_rank) == 0
                terminal = ws.basis_rank >= ws.basis_limit or ws.depth >= max_depth or \
                    not ws.can_advance()
                if (not feedback_pending and not matched_cost) or completes_pair or exhausts or \
                        terminal:
                    evaluate_checkpoint(not feedback_pending and not matched_cost)
                    res = st["result"]
                    if failed_numerically(res) or res.status == R.SUCCESS:
                        return finish(res)
                if feedback_pending and ws.depth >= fb_target:
                    feedback_pending = False
                    reliable = (not truncated and not broke and ws.basis_rank > start_rank and
                                ws.basis_rank > stage_start_rank and frontier == stage_target and
                                ws.active_width == stage_target)
                    recorded = reliable and C.record_matched_horizon_cost(
                        plan, frontier, ws.active_width, ws.depth - start_depth, op_s, loc_s)
                    exh = C.remaining_matched_horizon(plan, ws.basis_rank) == 0
                    ext = exh and C.matched_horizon_progress_stalled(
                        plan, active_certificate(st["result"], lam)) and \
                        C.extend_matched_horizon_to_basis_limit(plan, ws.basis_rank, ws.basis_limit)
                    if not recorded or (exh and not ext):
                        C.reset_matched_horizon_plan(plan)
                elif matched_cost:
                    reliable = (not truncated and not broke and ws.basis_rank > start_rank and
                                frontier == ws.active_width)
                    recorded = reliable and C.record_matched_horizon_cost(
                        plan, frontier, ws.active_width, ws.depth - start_depth, op_s, loc_s)
                    exh = C.remaining_matched_horizon(plan, ws.basis_rank) == 0
                    ext = exh and C.matched_horizon_progress_stalled(
                        plan, active_certificate(st["result"], lam)) and \
                        C.extend_matched_horizon_to_basis_limit(plan, ws.basis_rank, ws.basis_limit)
                    if not recorded or (exh and not ext):
                        C.reset_matched_horizon_plan(plan)
                elif not feedback_pending:
                    C.record_deepen_progress(hist, interval_cert,
                                             active_certificate(st["result"], lam), frontier,
                                             ws.active_width, ws.depth - start_depth, op_s, loc_s)
                continue
            if feedback_pending or matched_cost:
                feedback_pending = False
                C.reset_matched_horizon_plan(plan)
            continue

        reason = R.EXHAUSTED_SEARCH_SPACE
        if not feedback_pending and epoch >= max_epochs and st["aux"] < max_aux and avail_inc > 0:
            reason = R.MAXIMUM_EPOCHS
        return finish_at_limit(R.WORK_LIMIT, reason)


def solve_fixed_rbgk(op, b, regularization=0.0, tolerance=1e-8, failure_probability=1e-6,
                     auxiliary_width=4, maximum_depth=32, maximum_basis_size=0, seed=0, stream=0,
                     stop_early=True, clock=None) -> R.LeastSquaresResult:
    """fixed_rbgk.cpp:18-209."""
    clock = clock or WallClock()
    b = np.asarray(b, dtype=np.float64)
    validate_common_inputs(op, b, regularization, tolerance)
    if not math.isfinite(failure_probability) or not (0.0 < failure_probability < 1.0):
        raise ValueError("failure_probability must lie strictly between zero and one")
    if auxiliary_width < 0:
        raise ValueError("auxiliary_width must be nonnegative")
    if maximum_depth <= 0:
        raise ValueError("maximum_depth must be positive")
    if maximum_basis_size < 0:
        raise ValueError("maximum_basis_size must be nonnegative")
    total_start = clock.now()
    stats = R.RunStatistics()
    ws = Workspace(op, b, regularization, maximum_basis_size, stats, clock)
    aux = 0 if ws.bnorm == 0.0 else min(auxiliary_width, max(0, ws.basis_limit - 1))
    trace = []
    st = {"result": None, "depth": 0, "evaluated": -1}
    sched = CertificateCheckpointSchedule(regularization, tolerance)

    def evaluate():
        r = ws.evaluate(tolerance, 0, aux)
        sched.record_evaluation(st["depth"], r)
        st["evaluated"] = st["depth"]
        trace.append(r.trace[0])
        r.trace = []
        st["result"] = r

    def evaluate_terminal():
        if st["evaluated"] != st["depth"]:
            evaluate()
        return failed_numerically(st["result"])

    init = np.zeros((op.rows, aux))
    if aux > 0:
        init = gaussian_matrix(op.rows, aux, seed, stream, 0)
        stats.gaussian_random_block_requests += 1
        stats.gaussian_random_columns += aux
        stats.gaussian_random_values += op.rows * aux
    ws.initialize(init)
    if ws.can_advance():
        ws.advance_level()
    st["depth"] = ws.depth
    evaluate()
    truncated = step_has_numerical_rank_truncation(ws.audit)
    broke = step_has_nonfinite_breakdown(ws.audit)
    while True:
        res = st["result"]
        if failed_numerically(res):
            break
        if res.status == R.SUCCESS and stop_early:
            break
        if broke:
            if evaluate_terminal():
                break
            st["result"].status, st["result"].stop_reason = R.NUMERICAL_BREAKDOWN, \
                R.STOP_NUMERICAL_BREAKDOWN
            break
        terminal = None
        if ws.basis_rank >= ws.basis_limit:
            terminal = (R.BASIS_LIMIT, R.MAXIMUM_BASIS)
        elif st["depth"] >= maximum_depth:
            terminal = (R.WORK_LIMIT, R.MAXIMUM_DEPTH)
        elif truncated:
            terminal = (R.PRECISION_LIMIT, R.STOP_PRECISION_LIMIT)
        elif not ws.can_advance():
            terminal = (R.WORK_LIMIT, R.EXHAUSTED_SEARCH_SPACE)
        if terminal is not None:
            if evaluate_terminal():
                break
            if st["result"].status != R.SUCCESS:
                st["result"].status, st["result"].stop_reason = terminal
            break
        ok = ws.advance_level()
        truncated = step_has_numerical_rank_truncation(ws.audit)
        broke = step_has_nonfinite_breakdown(ws.audit)
        if not ok:
            if not ws.audit.right_rank_loss and not broke:
                raise RuntimeError("failed fixed extension did not report rank loss")
            continue
        st["depth"] = ws.depth
        term = st["depth"] >= maximum_depth or ws.basis_rank >= ws.basis_limit or \
            not ws.can_advance() or truncated or broke
        if sched.should_evaluate(st["depth"]) or term:
            evaluate()
    res = st["result"]
    res.iterations = st["depth"]
    res.depth = st["depth"]
    res.trace = trace
    stats.total_seconds = clock.now() - total_start
    res.statistics = R.RunStatistics(**vars(stats))
    return res


def solve_lsqr(op, b, regularization=0.0, tolerance=1e-8, maximum_iterations=0,
               clock=None) -> R.LeastSquaresResult:
    """lsqr.cpp:84-279."""
    clock = clock or WallClock()
    b = np.asarray(b, dtype=np.float64)
    validate_common_inputs(op, b, regularization, tolerance)
    if maximum_iterations < 0:
        raise ValueError("maximum_iterations must be nonnegative")
    max_it = maximum_iterations if maximum_iterations > 0 else min(op.rows, op.cols)
    total_start = clock.now()
    stats = R.RunStatistics()
    counted = Counted(op, stats, clock)

    def search(x, transpose):
        start = clock.now()
        clock.charge_operator(op.relative_block_product_cost(1))
        y = (op.apply_transpose if transpose else op.apply)(x.reshape(-1, 1)).ravel()
        sec = clock.now() - start
        if transpose:
            stats.at_seconds += sec
            stats.at_columns += 1
            stats.at_block_calls += 1
            stats.search_at_columns += 1
            stats.search_at_block_calls += 1
        else:
            stats.a_seconds += sec
            stats.a_columns += 1
            stats.a_block_calls += 1
            stats.search_a_columns += 1
            stats.search_a_block_calls += 1
        return np.asarray(y, dtype=np.float64)

    u = b.copy()
    beta = float(np.linalg.norm(u))
    x = np.zeros(op.cols)
    if beta > 0.0:
        u *= 1.0 / beta
    v = search(u, True)
    alpha = float(np.linalg.norm(v))
    if alpha > 0.0:
        v *= 1.0 / alpha
    w = v.copy()
    rhobar, phibar = alpha, beta
    damp = math.sqrt(regularization)
    trace = []
    sched = CertificateCheckpointSchedule(regularization, tolerance)
    st = {"result": None}

    def evaluate(it):
        r = validate_original_candidate(counted, b, x, regularization, tolerance)
        r.iterations = it
        r.depth = it
        r.basis_rank = it
        r.trace.append(R.make_iteration_record(r, 0, it, 0))
        sched.record_evaluation(it, r)
        trace.append(r.trace[0])
        r.trace = []
        st["result"] = r

    def done():
        res = st["result"]
        res.trace = trace
        stats.total_seconds = clock.now() - total_start
        res.statistics = R.RunStatistics(**vars(stats))
        return res

    if not math.isfinite(beta) or not math.isfinite(alpha) or alpha == 0.0:
        evaluate(0)
        res = st["result"]
        if failed_numerically(res) or res.status == R.SUCCESS or alpha == 0.0:
            return done()
    for it in range(1, max_it + 1):
        av = search(v, False)
        av -= alpha * u
        beta = float(np.linalg.norm(av))
        if beta > 0.0:
            av *= 1.0 / beta
        rhobar1 = math.hypot(rhobar, damp)
        cs1 = 1.0 if rhobar1 == 0.0 else rhobar / rhobar1
        phibar *= cs1
        rho = math.hypot(rhobar1, beta)
        cs = 1.0 if rho == 0.0 else rhobar1 / rho
        sn = 0.0 if rho == 0.0 else beta / rho
        phi = cs * phibar
        phibar = sn * phibar
        finite = all(math.isfinite(t) for t in (beta, rhobar1, cs1, rho, cs, sn, phi, phibar))
        if rho != 0.0:
            upd = phi / rho
            finite = finite and math.isfinite(upd)
            x += upd * w
        finite = finite and bool(np.all(np.isfinite(x)))
        u = av
        beta_breakdown = beta == 0.0
        checkpoint = sched.should_evaluate(it) or beta_breakdown or not finite or it == max_it
        evaluated = False
        if checkpoint:
            evaluate(it)
            evaluated = True
            res = st["result"]
            if failed_numerically(res) or res.status == R.SUCCESS:
                break
            if not finite:
                res.status, res.stop_reason = R.NUMERICAL_BREAKDOWN, R.STOP_NUMERICAL_BREAKDOWN
                break
            if beta_breakdown:
                res.status, res.stop_reason = R.WORK_LIMIT, R.EXHAUSTED_SEARCH_SPACE
                break
            if it == max_it:
                res.status, res.stop_reason = R.WORK_LIMIT, R.MAXIMUM_DEPTH
                break
        atu = search(u, True)
        atu -= beta * v
        next_alpha = float(np.linalg.norm(atu))
        cont = math.isfinite(next_alpha) and bool(np.all(np.isfinite(atu)))
        if next_alpha > 0.0 and math.isfinite(next_alpha):
            atu *= 1.0 / next_alpha
        theta = sn * next_alpha
        next_rhobar = -cs * next_alpha
        rec = theta / rho
        cont = cont and math.isfinite(theta) and math.isfinite(next_rhobar) and math.isfinite(rec)
        if cont and next_alpha > 0.0:
            w = atu - rec * w
            cont = bool(np.all(np.isfinite(w)))
        if not cont or next_alpha == 0.0:
            if not evaluated:
                evaluate(it)
                res = st["result"]
                if failed_numerically(res) or res.status == R.SUCCESS:
                    break
            res = st["result"]
            res.status = R.WORK_LIMIT if cont else R.NUMERICAL_BREAKDOWN
            res.stop_reason = R.EXHAUSTED_SEARCH_SPACE if cont else R.STOP_NUMERICAL_BREAKDOWN
            break
        alpha, rhobar = next_alpha, next_rhobar
        v = atu
    return done()
