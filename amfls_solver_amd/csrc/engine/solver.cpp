#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>
#include <limits>
#include <memory>
#include <stdexcept>

#include "controller.hpp"
#include "refine.hpp"
#include "validation.hpp"
#include "workspace.hpp"

namespace amflsx {
namespace {

int default_epochs(int dimension) {
    int epochs = 1, covered = 1;
    while (covered < dimension && epochs < 30) {
        covered *= 2;
        ++epochs;
    }
    return epochs;
}

double active_certificate(const LeastSquaresResult& r, double lambda) {
    return lambda > 0.0 ? r.relative_energy_error_upper_bound : r.backward_error_upper_bound;
}

void validate_common(const DeviceOperator& op, const double* b, double lambda, double tol) {
    if (op.rows() <= 0 || op.cols() <= 0) throw std::invalid_argument("operator dimensions must be positive");
    if (b == nullptr) throw std::invalid_argument("right-hand side cannot be null");
    if (!std::isfinite(lambda) || lambda < 0.0)
        throw std::invalid_argument("regularization must be finite and nonnegative");
    if (!std::isfinite(tol) || tol <= 0.0 || tol >= 1.0)
        throw std::invalid_argument("tolerance must lie strictly between zero and one");
}

void check_rhs_finite(Context& ctx, const double* b, int64_t m) {
    double st[3];
    ctx.block_stats<double>(m, 1, b, 1, st);
    if (st[2] > 0.0) throw std::invalid_argument("right-hand side must be finite");
}

// Recent-interval features for the learned width predictor (engine
// extension; the reference keeps only two same-width intervals).
struct IntervalLog {
    std::deque<double> decrease, seconds;
    void push(double prev, double cur, double sec_per_level) {
        double d = 0.0;
        if (prev > 0.0 && cur > 0.0 && std::isfinite(prev) && std::isfinite(cur)) d = std::log10(prev / cur);
        decrease.push_front(d);
        seconds.push_front(sec_per_level);
        if (decrease.size() > 4) decrease.pop_back();
        if (seconds.size() > 4) seconds.pop_back();
    }
    void clear() {
        decrease.clear();
        seconds.clear();
    }
};

double operator_nnz(const DeviceOperator& op) {
    if (auto* c = dynamic_cast<const CsrOperator*>(&op)) return static_cast<double>(c->nnz());
    return static_cast<double>(op.rows()) * static_cast<double>(op.cols());
}

// Evaluation of the driver's current basis at another regularization, and
// the observer the regularization sweep installs: called after every
// checkpoint evaluation (final = false) and once before the driver returns
// (final = true).
using LambdaEvaluator = std::function<LeastSquaresResult(double lambda)>;
using CheckpointHook = std::function<void(const LambdaEvaluator& evaluate_at, int depth, bool final)>;

// The adaptive driver (reference amfls.cpp:57-661) in working precision T.
// Called with the workspace just before the driver returns (the basis and
// the projected operator of the final state).
template <typename T>
using WorkspaceHook = std::function<void(Workspace<T>&)>;

template <typename T>
LeastSquaresResult amfls_driver(const DeviceOperator& op, const double* b_dev, const SolveOptions& o, Context& ctx,
                                const WidthPolicy* policy, const CheckpointHook* hook = nullptr,
                                const WorkspaceHook<T>* on_finish = nullptr) {
    const double lambda = o.regularization, tol = o.tolerance;
    Clock& clk = ctx.clock;
    const double total_start = clk.now();
    RunStatistics stats;
    ValidationEngine validator(op, b_dev, ctx, stats);
    Workspace<T> ws(op, b_dev, lambda, o.maximum_basis_size, stats, ctx, validator);
    ws.set_fp32_rank_factor(o.fp32_rank_factor);
    // Operator seconds so far (waits for the timed products in flight).
    auto op_seconds = [&] {
        ctx.timer.flush();
        return stats.a_seconds + stats.at_seconds;
    };
    const int dimension = static_cast<int>(std::min<int64_t>(op.rows(), op.cols()));
    const int max_depth = o.maximum_depth > 0 ? std::min(o.maximum_depth, dimension) : dimension;
    const int cap_aux = std::max(0, ws.basis_limit() - 1);
    const int max_aux = std::min(o.maximum_auxiliary_width > 0 ? o.maximum_auxiliary_width : cap_aux, cap_aux);
    const int max_epochs = o.maximum_epochs > 0 ? o.maximum_epochs : default_epochs(dimension);

    std::vector<IterationRecord> trace;
    ctl::CheckpointSchedule schedule(lambda, tol);
    int evaluations = 0, last_depth = -1, aux = 0;
    LeastSquaresResult result, best;
    double best_cert = kInfD;
    bool has = false;
    double last_op_s = 0.0, last_loc_s = 0.0;
    int last_interval_width = 0;

    auto record = [&](LeastSquaresResult cand, bool train) {
        if (cand.trace.size() != 1) throw std::logic_error("AMFLS candidate must contain exactly one trace record");
        if (cand.depth == last_depth) throw std::logic_error("AMFLS evaluated one checkpoint more than once");
        last_depth = cand.depth;
        if (train) schedule.record(cand.depth, cand);
        IterationRecord rec = cand.trace.front();
        rec.interval_operator_seconds = last_op_s;
        rec.interval_local_seconds = last_loc_s;
        rec.active_width = last_interval_width;
        trace.push_back(rec);
        cand.trace.clear();
        ++evaluations;
        const double cert = active_certificate(cand, lambda);
        const bool better = !has || (std::isfinite(cert) && !std::isfinite(best_cert)) || cert < best_cert;
        if (better) {
            best_cert = cert;
            best = cand;
        }
        has = true;
        result = std::move(cand);
    };
    // Inner solves of the refinement loop stop once the certificate has not
    // halved over the last doubling of the depth (an FP32 process has
    // reached its rounding floor; the outer FP64 loop continues from there).
    std::vector<std::pair<int, double>> eval_history;
    bool stagnated = false;
    const LambdaEvaluator evaluate_at = [&](double l) {
        const double keep = ws.regularization();
        ws.set_regularization(l);
        try {
            LeastSquaresResult r = ws.evaluate(tol, evaluations, aux);
            ws.set_regularization(keep);
            return r;
        } catch (...) {
            ws.set_regularization(keep);
            throw;
        }
    };
    auto evaluate = [&](bool train) {
        record(ws.evaluate(tol, evaluations, aux), train);
        if (hook) (*hook)(evaluate_at, ws.depth(), false);
        if (o.stop_on_stagnation) {
            const double c = active_certificate(result, lambda);
            const int d = result.depth;
            double earlier = kInfD;
            for (const auto& e : eval_history)
                if (2 * e.first <= d) earlier = std::min(earlier, e.second);
            eval_history.emplace_back(d, c);
            stagnated = d >= 16 && std::isfinite(earlier) && !(c < 0.5 * earlier);
        }
    };
    auto finish = [&](LeastSquaresResult f) {
        if (hook) (*hook)(evaluate_at, ws.depth(), true);
        if (on_finish) (*on_finish)(ws);
        f.iterations = ws.depth();
        f.trace = trace;
        ctx.timer.flush();
        // validation products timed in their kernels reach the parts on flush
        stats.validation_seconds = stats.base_validation_seconds + stats.ridge_correction_validation_seconds;
        stats.total_seconds = clk.now() - total_start;
        f.statistics = stats;
        return f;
    };
    auto finish_at_limit = [&](SolverStatus st, StopReason reason) {
        LeastSquaresResult f = lambda > 0.0 ? result : best;
        f.status = st;
        f.stop_reason = reason;
        return finish(std::move(f));
    };

    ws.initialize(false, 0, 0, 0, 0);
    if (ws.can_advance() && max_depth > 0) {
        last_interval_width = ws.active_left_width();
        const double t0 = clk.now();
        const double a0 = op_seconds();
        (void)ws.advance_level();
        last_op_s = op_seconds() - a0;
        last_loc_s = std::max(0.0, clk.now() - t0 - last_op_s);
    }
    evaluate(true);
    if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success)
        return finish(result);
    bool truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
    bool broke = step_has_nonfinite_breakdown(ws.last_step_audit());

    ctl::ProgressHistory hist;
    ctl::MatchedHorizonPlan plan;
    IntervalLog ilog;
    int epoch = 0;
    uint64_t random_generated = 0;
    bool feedback_pending = false;
    int fb_target = 0, stage_start_rank = 0, stage_origin = 0, stage_target = 0;
    const double nnz = operator_nnz(op);

    while (true) {
        if (stagnated) return finish_at_limit(SolverStatus::precision_limit, StopReason::precision_limit);
        if (broke) return finish_at_limit(SolverStatus::numerical_breakdown, StopReason::numerical_breakdown);
        if (ws.basis_rank() >= ws.basis_limit()) return finish_at_limit(SolverStatus::basis_limit, StopReason::maximum_basis);
        if (ws.depth() >= max_depth) return finish_at_limit(SolverStatus::work_limit, StopReason::maximum_depth);
        if (truncated) return finish_at_limit(SolverStatus::precision_limit, StopReason::precision_limit);

        const double interval_cert = active_certificate(result, lambda);
        const int frontier = ws.active_left_width();
        const int rem_basis = std::max(0, ws.basis_limit() - ws.basis_rank());
        const int avail_inc = std::max(0, rem_basis - frontier);
        const int rem_aux = std::max(0, max_aux - aux);
        const int width_cap = std::min(avail_inc, rem_aux);
        const bool fb_available = max_depth - ws.depth() >= 2;
        const int fb_depth = fb_available ? ws.depth() + 2 : ws.depth();
        const int levels_to_fb = fb_depth - ws.depth();
        const int fb_feasible = fb_available ? ctl::feedback_feasible_increment(ws.basis_rank(), ws.basis_limit(),
                                                                               frontier, levels_to_fb)
                                             : 0;
        const int feasible_cap = std::min(width_cap, fb_feasible);
        if (plan.end_rank >= 0 && ctl::remaining(plan, ws.basis_rank()) == 0) ctl::reset(plan);
        const int rem_horizon = ctl::remaining(plan, ws.basis_rank());
        const bool horizon_active = rem_horizon > 0;
        const bool collecting = horizon_active && !feedback_pending && plan.samples < 2;
        const bool ready = !feedback_pending && !collecting;
        const bool stage_available = ready && epoch < max_epochs;
        const int next_inc = stage_available ? ctl::next_dyadic_increment(frontier, aux, max_aux, feasible_cap) : 0;
        int requested = 0, requested_horizon = 0;
        if (stage_available && next_inc > 0) {
            const int target_active = frontier + next_inc;
            const ctl::WidthDecision d =
                horizon_active
                    ? ctl::horizon_candidate(next_inc, rem_horizon, ws.depth(), max_depth, ws.basis_rank(),
                                             ws.basis_limit(), frontier, plan.op_s, plan.loc_s, plan.samples,
                                             op.relative_block_product_cost(frontier),
                                             op.relative_block_product_cost(target_active), plan.origin_width > 1)
                    : ctl::initial_candidate(next_inc, interval_cert, tol, ws.depth(), max_depth, ws.basis_rank(),
                                             ws.basis_limit(), frontier, hist,
                                             op.relative_block_product_cost(frontier),
                                             op.relative_block_product_cost(target_active));
            bool widen = d.candidate;
            if (policy && *policy) {
                WidthFeatures f;
                f.active_width = frontier;
                f.proposed_increment = next_inc;
                for (size_t i = 0; i < ilog.decrease.size(); ++i) f.bound_decrease[i] = ilog.decrease[i];
                for (size_t i = 0; i < ilog.seconds.size(); ++i) f.step_seconds[i] = ilog.seconds[i];
                f.rows = static_cast<double>(op.rows());
                f.cols = static_cast<double>(op.cols());
                f.nnz = nnz;
                f.certificate = interval_cert;
                f.tolerance = tol;
                f.depth = ws.depth();
                f.basis_rank = ws.basis_rank();
                f.basis_limit = ws.basis_limit();
                f.heuristic = d.candidate;
                widen = (*policy)(f);
            }
            if (widen) {
                requested = next_inc;
                if (!horizon_active) requested_horizon = d.matched_horizon > 0 ? d.matched_horizon : 0;
            }
        }
        if (horizon_active && ready && requested == 0) ctl::reset(plan);

        if (requested > 0) {
            fb_target = fb_depth;
            stage_start_rank = ws.basis_rank();
            stage_origin = horizon_active ? plan.origin_width : frontier;
            stage_target = frontier + requested;
            ctl::reset(hist);
            ilog.clear();
            schedule = ctl::CheckpointSchedule(lambda, tol);
            stats.gaussian_random_block_requests += 1;
            stats.gaussian_random_columns += requested;
            stats.gaussian_random_values += op.rows() * static_cast<long long>(requested);
            const uint64_t first = random_generated;
            random_generated += static_cast<uint64_t>(requested);
            aux += requested;
            ++epoch;
            (void)ws.append_left_seed_gaussian(requested, o.seed, o.stream, first, true);
            bool advanced = false;
            if (ws.can_advance()) {
                last_interval_width = ws.active_left_width();
                const double t0 = clk.now();
                const double a0 = op_seconds();
                advanced = ws.advance_level();
                last_op_s = op_seconds() - a0;
                last_loc_s = std::max(0.0, clk.now() - t0 - last_op_s);
            }
            if (advanced) {
                if (requested_horizon > 0) {
                    (void)ctl::start_plan(plan, stage_start_rank, ws.basis_limit(), requested_horizon, stage_origin,
                                          stage_target, interval_cert);
                } else {
                    ctl::reset_costs(plan, stage_target);
                }
                truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
                broke = step_has_nonfinite_breakdown(ws.last_step_audit());
                const bool exhausted = ctl::remaining(plan, ws.basis_rank()) == 0;
                const bool terminal = ws.basis_rank() >= ws.basis_limit() || ws.depth() >= max_depth ||
                                      !ws.can_advance();
                if (terminal || exhausted) {
                    evaluate(false);
                    if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success)
                        return finish(result);
                }
                const bool reliable = !truncated && !broke && ws.basis_rank() > stage_start_rank &&
                                      ws.active_left_width() == stage_target;
                const bool extended = exhausted && ctl::progress_stalled(plan, active_certificate(result, lambda)) &&
                                      ctl::extend_to_limit(plan, ws.basis_rank(), ws.basis_limit());
                if (!reliable || (exhausted && !extended)) {
                    ctl::reset(plan);
                    feedback_pending = false;
                } else {
                    feedback_pending = true;
                }
            } else {
                truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
                broke = step_has_nonfinite_breakdown(ws.last_step_audit());
                feedback_pending = false;
                ctl::reset(plan);
            }
            continue;
        }

        if (ws.can_advance()) {
            const int start_depth = ws.depth();
            const int start_rank = ws.basis_rank();
            const bool matched_cost = !feedback_pending && ctl::remaining(plan, start_rank) > 0 && plan.samples < 2;
            int target_depth;
            if (feedback_pending) target_depth = fb_target;
            else if (matched_cost) target_depth = std::min(start_depth + 1, max_depth);
            else {
                target_depth = start_depth + 1;
                while (target_depth < max_depth && !schedule.should_evaluate(target_depth)) ++target_depth;
            }
            const double t0 = clk.now();
            const double a0 = op_seconds();
            last_interval_width = frontier;
            while (ws.can_advance() && ws.depth() < target_depth) {
                const bool advanced = ws.advance_toward(target_depth);
                truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
                broke = step_has_nonfinite_breakdown(ws.last_step_audit());
                if (!advanced || truncated || broke) break;
            }
            const double op_total = op_seconds();
            const double search_s = clk.now() - t0;
            const double op_s = std::max(0.0, op_total - a0);
            const double loc_s = std::max(0.0, search_s - op_s);
            last_op_s = op_s;
            last_loc_s = loc_s;
            if (ws.depth() > start_depth) {
                const bool completes_pair = matched_cost && plan.samples == 1;
                const bool exhausts = (feedback_pending || matched_cost) && ctl::remaining(plan, ws.basis_rank()) == 0;
                const bool terminal = ws.basis_rank() >= ws.basis_limit() || ws.depth() >= max_depth ||
                                      !ws.can_advance();
                if ((!feedback_pending && !matched_cost) || completes_pair || exhausts || terminal) {
                    evaluate(!feedback_pending && !matched_cost);
                    if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success)
                        return finish(result);
                }
                const int levels = ws.depth() - start_depth;
                if (feedback_pending && ws.depth() >= fb_target) {
                    feedback_pending = false;
                    const bool reliable = !truncated && !broke && ws.basis_rank() > start_rank &&
                                          ws.basis_rank() > stage_start_rank && frontier == stage_target &&
                                          ws.active_left_width() == stage_target;
                    const bool recorded = reliable && ctl::record_cost(plan, frontier, ws.active_left_width(), levels,
                                                                       op_s, loc_s);
                    const bool exh = ctl::remaining(plan, ws.basis_rank()) == 0;
                    const bool ext = exh && ctl::progress_stalled(plan, active_certificate(result, lambda)) &&
                                     ctl::extend_to_limit(plan, ws.basis_rank(), ws.basis_limit());
                    if (!recorded || (exh && !ext)) ctl::reset(plan);
                } else if (matched_cost) {
                    const bool reliable = !truncated && !broke && ws.basis_rank() > start_rank &&
                                          frontier == ws.active_left_width();
                    const bool recorded = reliable && ctl::record_cost(plan, frontier, ws.active_left_width(), levels,
                                                                       op_s, loc_s);
                    const bool exh = ctl::remaining(plan, ws.basis_rank()) == 0;
                    const bool ext = exh && ctl::progress_stalled(plan, active_certificate(result, lambda)) &&
                                     ctl::extend_to_limit(plan, ws.basis_rank(), ws.basis_limit());
                    if (!recorded || (exh && !ext)) ctl::reset(plan);
                } else if (!feedback_pending) {
                    const double cur = active_certificate(result, lambda);
                    ctl::record_deepen_progress(hist, interval_cert, cur, frontier, ws.active_left_width(), levels,
                                                op_s, loc_s);
                    if (frontier == ws.active_left_width())
                        ilog.push(interval_cert, cur, (op_s + loc_s) / std::max(1, levels));
                }
                continue;
            }
            if (feedback_pending || matched_cost) {
                feedback_pending = false;
                ctl::reset(plan);
            }
            continue;
        }
        StopReason reason = StopReason::exhausted_search_space;
        if (!feedback_pending && epoch >= max_epochs && aux < max_aux && avail_inc > 0)
            reason = StopReason::maximum_epochs;
        return finish_at_limit(SolverStatus::work_limit, reason);
    }
}

// Mixed precision: an FP32 block Golub-Kahan search (products and basis in
// FP32; projected solves and certificates in FP64) on the original problem,
// then FP64 LSQR refinement preconditioned by the search's factorization
// (refine.hpp): every vector that carries the residual is FP64, so the
// refinement reaches the FP64 solution instead of stalling at the FP32
// search's floor.
LeastSquaresResult mixed_driver(const DeviceOperator& op, const double* b_dev, const SolveOptions& o, Context& ctx,
                                const WidthPolicy* policy) {
    const int64_t n = op.cols();
    Clock& clk = ctx.clock;
    const double total_start = clk.now();
    BasisCapture basis;
    const WorkspaceHook<float> capture = [&](Workspace<float>& ws) {
        basis.n = n;
        basis.k = ws.basis_rank();
        basis.T = ws.projected_operator();
        if (basis.k > 0) {
            basis.V.reset(static_cast<size_t>(n) * static_cast<size_t>(basis.k));
            copy_block<float, float>(ctx.stream, n, basis.k, ws.basis(), ws.basis_ld(), basis.V.get(), basis.k, 1.0);
        }
        basis.valid = basis.k > 0 && basis.T.cols() == basis.k;
    };
    SolveOptions so = o;
    so.precision = Precision::fp32;
    // The search stops once its FP64-validated certificate stagnates (its
    // candidates are FP32-accurate at best); the refinement takes it from
    // there.  A forward-error target on an unregularized problem with at
    // least as many rows as columns needs every direction of R^n resolved,
    // so the search then runs on to a complete basis (or a rank truncation)
    // and the preconditioner covers all of R^n.
    const bool full_basis = o.regularization == 0.0 && o.forward_tolerance > 0.0 && op.rows() >= op.cols();
    so.stop_on_stagnation = !full_basis;
    // The search's rank rule only has to reject rounding noise: after CGS2
    // an FP32 residual block carries about eps32 * scale of it.  Directions
    // kept above that are real (a noisy one only weakens the preconditioner,
    // which the FP64 refinement absorbs), so the cutoff is 8 eps32 * scale
    // rather than the FP32 solver's 64 eps32 sqrt(k) * scale, which on the
    // 1e6 x 2000 dense problem truncates the basis at rank 1730 of 2000.
    if (so.fp32_rank_factor <= 0.0) so.fp32_rank_factor = 8.0 * std::numeric_limits<float>::epsilon();
    if (o.inner_tolerance > 0.0) so.tolerance = o.inner_tolerance;
    LeastSquaresResult search = amfls_driver<float>(op, b_dev, so, ctx, policy, nullptr, &capture);
    const bool forward = o.forward_tolerance > 0.0;
    const bool settled = search.status == SolverStatus::success && so.tolerance <= o.tolerance && !forward;
    if (!basis.valid || o.maximum_refinements == 0 || settled || candidate_validation_failed_numerically(search) ||
        search.solution.size() != static_cast<size_t>(n)) {
        search.statistics.total_seconds = clk.now() - total_start;
        return search;
    }
    RunStatistics stats = search.statistics;
    DevBuf<double> x(static_cast<size_t>(n));
    AMFLS_HIP_CHECK(hipMemcpyAsync(x.get(), search.solution.data(), static_cast<size_t>(n) * sizeof(double),
                                   hipMemcpyHostToDevice, ctx.stream));
    std::vector<IterationRecord> trace = search.trace;
    RefineReport rep;
    LeastSquaresResult r = refine_preconditioned(op, b_dev, o, ctx, stats, basis, x.get(), trace, rep);
    const bool target_met = std::isfinite(r.forward_error_estimate) && r.forward_error_estimate <= o.forward_tolerance;
    if (forward && !target_met) {
        // The search's factorization did not precondition the problem well
        // enough for the refinement to reach the forward-error target within
        // its budget (a search stopped early by a rank truncation leaves part
        // of R^n to LSQR): the FP64 engine solves it instead.  Its result is
        // returned, with the trace and the product counters of both runs.  (A
        // met target with a certificate stuck at its rounding floor is kept:
        // the FP64 engine stops at the same floor.)
        SolveOptions fo = o;
        fo.precision = Precision::fp64;
        LeastSquaresResult f = amfls_driver<double>(op, b_dev, fo, ctx, policy);
        RunStatistics fs = f.statistics;
        fs.a_columns += stats.a_columns;
        fs.at_columns += stats.at_columns;
        fs.a_block_calls += stats.a_block_calls;
        fs.at_block_calls += stats.at_block_calls;
        fs.a_seconds += stats.a_seconds;
        fs.at_seconds += stats.at_seconds;
        fs.iterative_a_columns += stats.iterative_a_columns;
        fs.iterative_at_columns += stats.iterative_at_columns;
        fs.iterative_a_block_calls += stats.iterative_a_block_calls;
        fs.iterative_at_block_calls += stats.iterative_at_block_calls;
        trace.insert(trace.end(), f.trace.begin(), f.trace.end());
        f.trace = std::move(trace);
        f.refinement_steps = rep.iterations;
        f.iterations += search.iterations + rep.iterations;
        fs.total_seconds = clk.now() - total_start;
        f.statistics = fs;
        f.fp64_fallback = true;
        return f;
    }
    if (search.status == SolverStatus::success && r.status != SolverStatus::success && !forward) r = search;
    r.depth = search.depth;
    r.auxiliary_width = search.auxiliary_width;
    r.basis_rank = search.basis_rank;
    r.refinement_steps = rep.iterations;
    r.iterations = search.iterations + rep.iterations;
    r.trace = std::move(trace);
    ctx.timer.flush();
    stats.validation_seconds = stats.base_validation_seconds + stats.ridge_correction_validation_seconds;
    stats.total_seconds = clk.now() - total_start;
    r.statistics = stats;
    return r;
}

}  // namespace

LeastSquaresResult solve_amfls(const DeviceOperator& op, const double* b_dev, const SolveOptions& o, Context& ctx,
                               const WidthPolicy* policy) {
    validate_common(op, b_dev, o.regularization, o.tolerance);
    if (!std::isfinite(o.failure_probability) || o.failure_probability <= 0.0 || o.failure_probability >= 1.0)
        throw std::invalid_argument("failure_probability must lie strictly between zero and one");
    if (o.maximum_epochs < 0 || o.maximum_depth < 0 || o.maximum_auxiliary_width < 0 || o.maximum_basis_size < 0)
        throw std::invalid_argument("work limits must be nonnegative");
    check_rhs_finite(ctx, b_dev, op.rows());
    ctx.clock.configure(o.clock_mode, o.synthetic_operator_unit, o.synthetic_local_unit);
    if (o.precision == Precision::mixed) return mixed_driver(op, b_dev, o, ctx, policy);
    if (o.precision == Precision::fp32) return amfls_driver<float>(op, b_dev, o, ctx, policy);
    return amfls_driver<double>(op, b_dev, o, ctx, policy);
}

namespace {

void validate_amfls_options(const SolveOptions& o) {
    if (!std::isfinite(o.failure_probability) || o.failure_probability <= 0.0 || o.failure_probability >= 1.0)
        throw std::invalid_argument("failure_probability must lie strictly between zero and one");
    if (o.maximum_epochs < 0 || o.maximum_depth < 0 || o.maximum_auxiliary_width < 0 || o.maximum_basis_size < 0)
        throw std::invalid_argument("work limits must be nonnegative");
}

}  // namespace

std::vector<LeastSquaresResult> solve_amfls_sweep(const DeviceOperator& op, const double* b_dev,
                                                  const SolveOptions& o, const std::vector<double>& lambdas,
                                                  Context& ctx, const WidthPolicy* policy) {
    if (lambdas.empty()) throw std::invalid_argument("the regularization sweep needs at least one value");
    for (double l : lambdas) validate_common(op, b_dev, l, o.tolerance);
    validate_amfls_options(o);
    if (o.precision == Precision::mixed)
        throw std::invalid_argument("the regularization sweep runs in fp64 or fp32 precision");
    check_rhs_finite(ctx, b_dev, op.rows());
    ctx.clock.configure(o.clock_mode, o.synthetic_operator_unit, o.synthetic_local_unit);
    const size_t L = lambdas.size();
    size_t drive = 0;
    for (size_t j = 1; j < L; ++j)
        if (lambdas[j] < lambdas[drive]) drive = j;
    const double lam_d = lambdas[drive];
    struct Slot {
        bool done = false, has = false;
        int evaluated_depth = -1;
        double best_cert = kInfD;
        LeastSquaresResult best;
        std::vector<IterationRecord> trace;
    };
    std::vector<Slot> slot(L);
    const CheckpointHook hook = [&](const LambdaEvaluator& evaluate_at, int depth, bool) {
        for (size_t j = 0; j < L; ++j) {
            Slot& sl = slot[j];
            if (sl.done || lambdas[j] == lam_d || depth == sl.evaluated_depth) continue;
            LeastSquaresResult r = evaluate_at(lambdas[j]);
            sl.evaluated_depth = depth;
            if (!r.trace.empty()) sl.trace.push_back(r.trace.front());
            r.trace.clear();
            const double c = active_certificate(r, lambdas[j]);
            if (!sl.has || (std::isfinite(c) && !std::isfinite(sl.best_cert)) || c < sl.best_cert) {
                sl.best_cert = c;
                sl.best = r;
                sl.has = true;
            }
            if (r.status == SolverStatus::success || candidate_validation_failed_numerically(r)) {
                sl.done = true;
                sl.best = std::move(r);
            }
        }
    };
    SolveOptions od = o;
    od.regularization = lam_d;
    LeastSquaresResult rd = o.precision == Precision::fp32 ? amfls_driver<float>(op, b_dev, od, ctx, policy, &hook)
                                                           : amfls_driver<double>(op, b_dev, od, ctx, policy, &hook);
    std::vector<LeastSquaresResult> out(L);
    for (size_t j = 0; j < L; ++j) {
        if (lambdas[j] == lam_d) {
            out[j] = rd;
            continue;
        }
        Slot& sl = slot[j];
        if (sl.done && sl.best.status == SolverStatus::success) {
            out[j] = std::move(sl.best);
            out[j].trace = std::move(sl.trace);
            out[j].iterations = out[j].depth;
            out[j].statistics = rd.statistics;
            continue;
        }
        // not certified along the shared process: a solve of its own
        SolveOptions oj = o;
        oj.regularization = lambdas[j];
        out[j] = o.precision == Precision::fp32 ? amfls_driver<float>(op, b_dev, oj, ctx, policy)
                                                : amfls_driver<double>(op, b_dev, oj, ctx, policy);
    }
    return out;
}

std::vector<std::vector<LeastSquaresResult>> solve_amfls_batch(const std::vector<BatchProblem>& problems,
                                                               const SolveOptions& o, int workers,
                                                               const WidthPolicy* policy) {
    for (const BatchProblem& p : problems)
        if (p.op == nullptr || p.b_dev == nullptr) throw std::invalid_argument("batch problem without operator or b");
    if (workers <= 0) throw std::invalid_argument("the batch needs at least one worker");
    std::vector<std::vector<LeastSquaresResult>> out(problems.size());
    if (problems.empty()) return out;
    workers = static_cast<int>(std::min<size_t>(static_cast<size_t>(workers), problems.size()));
    int device = 0;
    AMFLS_HIP_CHECK(hipGetDevice(&device));
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr error;
    std::mutex error_mu, policy_mu;
    // a stateful policy (the learned model) is called by one worker at a time
    WidthPolicy serial;
    if (policy && *policy)
        serial = [&](const WidthFeatures& f) {
            std::lock_guard<std::mutex> g(policy_mu);
            return (*policy)(f);
        };
    auto work = [&] {
        try {
            AMFLS_HIP_CHECK(hipSetDevice(device));
            Context wctx;
            while (!failed.load()) {
                const size_t i = next.fetch_add(1);
                if (i >= problems.size()) break;
                const BatchProblem& p = problems[i];
                out[i] = solve_amfls_sweep(*p.op, p.b_dev, o, p.lambdas, wctx, serial ? &serial : nullptr);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(error_mu);
            if (!error) error = std::current_exception();
            failed.store(true);
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(workers));
    for (int w = 0; w < workers; ++w) pool.emplace_back(work);
    for (std::thread& t : pool) t.join();
    if (error) std::rethrow_exception(error);
    return out;
}

LeastSquaresResult solve_fixed_rbgk(const DeviceOperator& op, const double* b_dev, const FixedRbgkOptions& o,
                                    Context& ctx) {
    // reference fixed_rbgk.cpp:18-209
    validate_common(op, b_dev, o.regularization, o.tolerance);
    if (!std::isfinite(o.failure_probability) || o.failure_probability <= 0.0 || o.failure_probability >= 1.0)
        throw std::invalid_argument("failure_probability must lie strictly between zero and one");
    if (o.auxiliary_width < 0) throw std::invalid_argument("auxiliary_width must be nonnegative");
    if (o.maximum_depth <= 0) throw std::invalid_argument("maximum_depth must be positive");
    if (o.maximum_basis_size < 0) throw std::invalid_argument("maximum_basis_size must be nonnegative");
    check_rhs_finite(ctx, b_dev, op.rows());
    ctx.clock.configure(o.clock_mode, 1e-3, 1e-6);
    Clock& clk = ctx.clock;
    const double total_start = clk.now();
    RunStatistics stats;
    ValidationEngine validator(op, b_dev, ctx, stats);
    Workspace<double> ws(op, b_dev, o.regularization, o.maximum_basis_size, stats, ctx, validator);
    const int aux = ws.right_hand_side_norm() == 0.0 ? 0 : std::min(o.auxiliary_width, std::max(0, ws.basis_limit() - 1));
    std::vector<IterationRecord> trace;
    LeastSquaresResult result;
    int depth = 0, evaluated = -1;
    ctl::CheckpointSchedule schedule(o.regularization, o.tolerance);
    auto evaluate = [&] {
        result = ws.evaluate(o.tolerance, 0, aux);
        schedule.record(depth, result);
        evaluated = depth;
        trace.push_back(result.trace.front());
        result.trace.clear();
    };
    auto evaluate_terminal = [&] {
        if (evaluated != depth) evaluate();
        return candidate_validation_failed_numerically(result);
    };
    if (aux > 0) {
        stats.gaussian_random_block_requests += 1;
        stats.gaussian_random_columns += aux;
        stats.gaussian_random_values += op.rows() * static_cast<long long>(aux);
    }
    ws.initialize(aux > 0, aux, o.seed, o.stream, 0);
    if (ws.can_advance()) (void)ws.advance_level();
    depth = ws.depth();
    evaluate();
    bool truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
    bool broke = step_has_nonfinite_breakdown(ws.last_step_audit());
    while (true) {
        if (candidate_validation_failed_numerically(result)) break;
        if (result.status == SolverStatus::success && o.stop_early) break;
        if (broke) {
            if (evaluate_terminal()) break;
            result.status = SolverStatus::numerical_breakdown;
            result.stop_reason = StopReason::numerical_breakdown;
            break;
        }
        SolverStatus ts = SolverStatus::success;
        StopReason tr = StopReason::certified_optimality;
        bool terminal = true;
        if (ws.basis_rank() >= ws.basis_limit()) { ts = SolverStatus::basis_limit; tr = StopReason::maximum_basis; }
        else if (depth >= o.maximum_depth) { ts = SolverStatus::work_limit; tr = StopReason::maximum_depth; }
        else if (truncated) { ts = SolverStatus::precision_limit; tr = StopReason::precision_limit; }
        else if (!ws.can_advance()) { ts = SolverStatus::work_limit; tr = StopReason::exhausted_search_space; }
        else terminal = false;
        if (terminal) {
            if (evaluate_terminal()) break;
            if (result.status != SolverStatus::success) {
                result.status = ts;
                result.stop_reason = tr;
            }
            break;
        }
        const bool ok = ws.advance_level();
        truncated = step_has_numerical_rank_truncation(ws.last_step_audit());
        broke = step_has_nonfinite_breakdown(ws.last_step_audit());
        if (!ok) {
            if (!ws.last_step_audit().right_rank_loss && !broke)
                throw std::logic_error("failed fixed extension did not report rank loss");
            continue;
        }
        depth = ws.depth();
        const bool term = depth >= o.maximum_depth || ws.basis_rank() >= ws.basis_limit() || !ws.can_advance() ||
                          truncated || broke;
        if (schedule.should_evaluate(depth) || term) evaluate();
    }
    result.iterations = depth;
    result.depth = depth;
    result.trace = trace;
    ctx.timer.flush();
    stats.validation_seconds = stats.base_validation_seconds + stats.ridge_correction_validation_seconds;
    stats.total_seconds = clk.now() - total_start;
    result.statistics = stats;
    return result;
}

LeastSquaresResult solve_lsqr(const DeviceOperator& op, const double* b_dev, const LsqrOptions& o, Context& ctx) {
    // reference lsqr.cpp:84-279 on device vectors (FP64)
    validate_common(op, b_dev, o.regularization, o.tolerance);
    if (o.maximum_iterations < 0) throw std::invalid_argument("maximum_iterations must be nonnegative");
    check_rhs_finite(ctx, b_dev, op.rows());
    ctx.clock.configure(o.clock_mode, 1e-3, 1e-6);
    Clock& clk = ctx.clock;
    const int64_t m = op.rows(), n = op.cols();
    const int max_it = o.maximum_iterations > 0 ? o.maximum_iterations : static_cast<int>(std::min(m, n));
    const double total_start = clk.now();
    RunStatistics stats;
    ValidationEngine validator(op, b_dev, ctx, stats);
    hipStream_t s = ctx.stream;
    DevBuf<double> u(m), v(n), w(n), x(n), av(m), atu(n);
    auto nrm = [&](const double* p, int64_t len) {
        double st[3];
        ctx.block_stats<double>(len, 1, p, 1, st);
        return fp::nominal_norm(st[0], st[1], st[2]);
    };
    auto search = [&](const double* in, double* out, bool transpose) {
        if (!clk.synthetic()) ctx.sync();
        const double start = clk.now();
        clk.charge_operator(op.relative_block_product_cost(1));
        if (transpose) op.apply_t(s, in, 1, 1, out, 1);
        else op.apply(s, in, 1, 1, out, 1);
        if (!clk.synthetic()) ctx.sync();
        const double sec = clk.now() - start;
        if (transpose) {
            stats.at_seconds += sec; ++stats.at_columns; ++stats.at_block_calls; ++stats.search_at_columns;
            ++stats.search_at_block_calls;
        } else {
            stats.a_seconds += sec; ++stats.a_columns; ++stats.a_block_calls; ++stats.search_a_columns;
            ++stats.search_a_block_calls;
        }
    };
    copy_block<double, double>(s, m, 1, b_dev, 1, u.get(), 1, 1.0);
    fill<double>(s, x.get(), n, 0.0);
    double beta = nrm(u.get(), m);
    if (beta > 0.0) scale_vec(s, m, 1.0 / beta, u.get(), u.get());
    search(u.get(), v.get(), true);
    double alpha = nrm(v.get(), n);
    if (alpha > 0.0) scale_vec(s, n, 1.0 / alpha, v.get(), v.get());
    copy_block<double, double>(s, n, 1, v.get(), 1, w.get(), 1, 1.0);
    double rhobar = alpha, phibar = beta;
    const double damp = std::sqrt(o.regularization);
    std::vector<IterationRecord> trace;
    LeastSquaresResult result;
    ctl::CheckpointSchedule schedule(o.regularization, o.tolerance);
    auto evaluate = [&](int it) {
        result = validator.validate(x.get(), o.regularization, o.tolerance);
        result.iterations = result.depth = result.basis_rank = it;
        result.trace.push_back(make_iteration_record(result, 0, it, 0));
        schedule.record(it, result);
        trace.push_back(result.trace.front());
        result.trace.clear();
    };
    auto done = [&] {
        result.trace = trace;
        stats.total_seconds = clk.now() - total_start;
        result.statistics = stats;
        return result;
    };
    if (!std::isfinite(beta) || !std::isfinite(alpha) || alpha == 0.0) {
        evaluate(0);
        if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success || alpha == 0.0)
            return done();
    }
    for (int it = 1; it <= max_it; ++it) {
        search(v.get(), av.get(), false);
        axpy_block<double>(s, m, 1, -alpha, u.get(), 1, av.get(), 1);
        beta = nrm(av.get(), m);
        if (beta > 0.0) scale_vec(s, m, 1.0 / beta, av.get(), av.get());
        const double rhobar1 = std::hypot(rhobar, damp);
        const double cs1 = rhobar1 == 0.0 ? 1.0 : rhobar / rhobar1;
        phibar *= cs1;
        const double rho = std::hypot(rhobar1, beta);
        const double cs = rho == 0.0 ? 1.0 : rhobar1 / rho;
        const double sn = rho == 0.0 ? 0.0 : beta / rho;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        bool finite = std::isfinite(beta) && std::isfinite(rhobar1) && std::isfinite(cs1) && std::isfinite(rho) &&
                      std::isfinite(cs) && std::isfinite(sn) && std::isfinite(phi) && std::isfinite(phibar);
        if (rho != 0.0) {
            const double upd = phi / rho;
            finite = finite && std::isfinite(upd);
            axpy_block<double>(s, n, 1, upd, w.get(), 1, x.get(), 1);
        }
        finite = finite && std::isfinite(nrm(x.get(), n));
        std::swap(u, av);
        const bool beta_breakdown = beta == 0.0;
        const bool checkpoint = schedule.should_evaluate(it) || beta_breakdown || !finite || it == max_it;
        bool evaluated = false;
        if (checkpoint) {
            evaluate(it);
            evaluated = true;
            if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success) break;
            if (!finite) { result.status = SolverStatus::numerical_breakdown; result.stop_reason = StopReason::numerical_breakdown; break; }
            if (beta_breakdown) { result.status = SolverStatus::work_limit; result.stop_reason = StopReason::exhausted_search_space; break; }
            if (it == max_it) { result.status = SolverStatus::work_limit; result.stop_reason = StopReason::maximum_depth; break; }
        }
        search(u.get(), atu.get(), true);
        axpy_block<double>(s, n, 1, -beta, v.get(), 1, atu.get(), 1);
        const double next_alpha = nrm(atu.get(), n);
        bool cont = std::isfinite(next_alpha);
        if (next_alpha > 0.0 && std::isfinite(next_alpha)) scale_vec(s, n, 1.0 / next_alpha, atu.get(), atu.get());
        const double theta = sn * next_alpha;
        const double next_rhobar = -cs * next_alpha;
        const double rec = theta / rho;
        cont = cont && std::isfinite(theta) && std::isfinite(next_rhobar) && std::isfinite(rec);
        if (cont && next_alpha > 0.0) {
            // w = atu - rec * w
            copy_block<double, double>(s, n, 1, w.get(), 1, w.get(), 1, -rec);
            axpy_block<double>(s, n, 1, 1.0, atu.get(), 1, w.get(), 1);
            cont = std::isfinite(nrm(w.get(), n));
        }
        if (!cont || next_alpha == 0.0) {
            if (!evaluated) {
                evaluate(it);
                if (candidate_validation_failed_numerically(result) || result.status == SolverStatus::success) break;
            }
            result.status = cont ? SolverStatus::work_limit : SolverStatus::numerical_breakdown;
            result.stop_reason = cont ? StopReason::exhausted_search_space : StopReason::numerical_breakdown;
            break;
        }
        alpha = next_alpha;
        rhobar = next_rhobar;
        std::swap(v, atu);
    }
    return done();
}

}  // namespace amflsx
