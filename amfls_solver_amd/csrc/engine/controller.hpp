// Width-controller helpers of the adaptive AMFLS driver.
//
// Restates the reference's pure decision functions
// (src/algorithms/amfls/amfls_internal.hpp:14-674): same-width progress
// history, the dyadic width increment, the matched-horizon cost gate and plan
// bookkeeping.  The decision rules are kept identical because the parity
// tests replay them against the CPU restatement with a synthetic clock.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>

namespace amflsx::ctl {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr int kIntMax = std::numeric_limits<int>::max();
constexpr double kDblMax = std::numeric_limits<double>::max();
using ld = long double;

struct ProgressHistory {
    double lc[2] = {0.0, 0.0};        // interval log contractions
    double level_lc[2] = {0.0, 0.0};  // per-level log contractions
    int levels[2] = {0, 0};
    double op_s[2] = {kInf, kInf};
    double loc_s[2] = {kInf, kInf};
    int count = 0;
    double prior_lc = 0.0;
    int prior_levels = 0;
    int active_width = -1;
};

inline void reset(ProgressHistory& h) { h = ProgressHistory{}; }

inline void record_deepen_progress(ProgressHistory& h, double prev, double cur, int w0, int w1, int levels,
                                   double op_s, double loc_s) {
    if (w0 <= 0 || w1 <= 0 || w0 != w1) {
        reset(h);
        return;
    }
    if (h.active_width != w0) {
        reset(h);
        h.active_width = w0;
    }
    if (levels <= 0 || !(prev > 0.0) || !(cur > 0.0) || !std::isfinite(prev) || !std::isfinite(cur)) return;
    const ld progress = prev > cur ? std::log(static_cast<ld>(prev)) - std::log(static_cast<ld>(cur)) : 0.0L;
    if (!(progress >= 0.0L) || !std::isfinite(progress)) return;
    const ld per_level = progress / static_cast<ld>(levels);
    const double per_level_d = per_level > 0.0L && std::isfinite(per_level)
                                   ? static_cast<double>(std::min(per_level, static_cast<ld>(kDblMax)))
                                   : 0.0;
    const bool finite_cost = op_s >= 0.0 && loc_s >= 0.0 && std::isfinite(op_s) && std::isfinite(loc_s) &&
                             op_s + loc_s > 0.0 && std::isfinite(op_s + loc_s);
    const double lop = finite_cost ? op_s / levels : kInf;
    const double lloc = finite_cost ? loc_s / levels : kInf;
    if (h.count < 2) {
        const int i = h.count++;
        h.lc[i] = static_cast<double>(progress);
        h.level_lc[i] = per_level_d;
        h.levels[i] = levels;
        h.op_s[i] = lop;
        h.loc_s[i] = lloc;
        return;
    }
    const ld pooled = static_cast<ld>(h.prior_lc) + static_cast<ld>(h.lc[0]);
    if (h.prior_levels > kIntMax - h.levels[0] || !std::isfinite(pooled) || pooled > static_cast<ld>(kDblMax)) {
        h.prior_lc = kInf;
        h.prior_levels = kIntMax;
    } else {
        h.prior_lc = static_cast<double>(pooled);
        h.prior_levels += h.levels[0];
    }
    h.lc[0] = h.lc[1];
    h.lc[1] = static_cast<double>(progress);
    h.level_lc[0] = h.level_lc[1];
    h.level_lc[1] = per_level_d;
    h.levels[0] = h.levels[1];
    h.levels[1] = levels;
    h.op_s[0] = h.op_s[1];
    h.op_s[1] = lop;
    h.loc_s[0] = h.loc_s[1];
    h.loc_s[1] = lloc;
}

inline int next_dyadic_increment(int active, int generated, int max_generated, int available) {
    if (active <= 0 || generated < 0 || max_generated <= generated || active > generated + 1 || available <= 0)
        return 0;
    const long long max_active = static_cast<long long>(active) + max_generated - generated;
    const long long cur = active;
    const long long next = cur > max_active / 2 ? max_active : std::min(2 * cur, max_active);
    if (next <= cur) return 0;
    const long long req = next - cur;
    return req <= available && req <= kIntMax ? static_cast<int>(req) : 0;
}

inline int ceiling_ratio(int a, int b) { return (a <= 0 || b <= 0) ? 0 : 1 + (a - 1) / b; }

inline int feedback_feasible_increment(int rank, int limit, int frontier, int levels) {
    if (rank < 0 || limit < rank || frontier < 0 || levels <= 0) return 0;
    return std::max(0, (limit - rank) / levels - frontier);
}

struct WidthDecision {
    double recent_level_lc = 0.0, prior_level_lc = 0.0, remaining_lc = 0.0;
    int forecast_levels = 0, incumbent_levels = 0, target_levels = 0, target_remaining_levels = 0;
    int matched_horizon = 0;
    double incumbent_seconds = kInf, target_seconds = kInf;
    bool history_mature = false, prior_history_mature = false, recent_slowdown = false,
         persistent_slowdown = false, progress_stalled = false, long_horizon_pressure = false,
         widening_necessary = false, observed_costs_mature = false, width_depth_admissible = true,
         pass_reduction_admissible = false, stage_feasible = false, finite_current_certificate = false,
         finite_resource_boundary = false, candidate = false;
};

inline void cost_gate(WidthDecision& d, int active, int target, const double* op_s, const double* loc_s, int samples,
                      double inc_cost, double tgt_cost) {
    if (d.incumbent_levels <= 0 || d.target_levels <= 0 || active <= 0 || target <= active || samples != 2 ||
        !(inc_cost > 0.0) || !(tgt_cost > 0.0) || !std::isfinite(inc_cost) || !std::isfinite(tgt_cost))
        return;
    const ld ratio = static_cast<ld>(target) / static_cast<ld>(active);
    const ld ratio2 = ratio * ratio;
    const ld op_ratio = static_cast<ld>(tgt_cost) / static_cast<ld>(inc_cost);
    const ld inc_op = static_cast<ld>(d.incumbent_levels) * static_cast<ld>(inc_cost);
    const ld tgt_op = static_cast<ld>(d.target_levels) * static_cast<ld>(tgt_cost);
    ld best_inc = std::numeric_limits<ld>::infinity(), worst_tgt = 0.0L;
    bool finite = true;
    for (int i = 0; i < samples; ++i) {
        const ld o = op_s[i], l = loc_s[i];
        if (!(o >= 0.0L) || !(l >= 0.0L) || !std::isfinite(o) || !std::isfinite(l) || !(o + l > 0.0L)) {
            finite = false;
            break;
        }
        const ld inc_s = static_cast<ld>(d.incumbent_levels) * (o + l);
        const ld tloc = l * ratio2;
        const ld tgt_s = static_cast<ld>(d.target_levels) * (o * op_ratio + tloc) + tloc;
        if (!std::isfinite(inc_s) || !std::isfinite(tgt_s)) {
            finite = false;
            break;
        }
        best_inc = std::min(best_inc, inc_s);
        worst_tgt = std::max(worst_tgt, tgt_s);
    }
    if (finite) {
        finite = best_inc <= static_cast<ld>(kDblMax) && worst_tgt <= static_cast<ld>(kDblMax);
        if (finite) {
            d.incumbent_seconds = static_cast<double>(best_inc);
            d.target_seconds = static_cast<double>(worst_tgt);
        }
    }
    d.observed_costs_mature = finite;
    d.pass_reduction_admissible = d.width_depth_admissible && d.target_levels < d.incumbent_levels &&
                                  tgt_op < inc_op && d.observed_costs_mature && d.target_seconds < d.incumbent_seconds;
}

inline WidthDecision horizon_candidate(int inc, int remaining, int depth, int max_depth, int rank, int limit, int active,
                                       const double* op_s, const double* loc_s, int samples, double inc_cost,
                                       double tgt_cost, bool require_balance) {
    WidthDecision d;
    d.stage_feasible = inc > 0 && active > 0 && inc <= kIntMax - active;
    d.finite_resource_boundary = depth >= 0 && max_depth >= depth && rank >= 0 && limit >= rank && active > 0 &&
                                 remaining > 0 && remaining <= limit - rank;
    if (!d.stage_feasible || !d.finite_resource_boundary) return d;
    const int target = active + inc;
    d.target_remaining_levels = std::min(max_depth - depth, (limit - rank) / target);
    const long long cap = static_cast<long long>(d.target_remaining_levels) * target;
    d.matched_horizon = static_cast<int>(std::min<long long>(remaining, cap));
    d.incumbent_levels = ceiling_ratio(d.matched_horizon, active);
    d.target_levels = ceiling_ratio(d.matched_horizon, target);
    if (d.target_levels < 2) return d;
    d.width_depth_admissible = !require_balance || d.target_levels > target;
    cost_gate(d, active, target, op_s, loc_s, samples, inc_cost, tgt_cost);
    d.candidate = d.pass_reduction_admissible;
    return d;
}

inline WidthDecision initial_candidate(int inc, double cert, double tol, int depth, int max_depth, int rank, int limit,
                                       int active, const ProgressHistory& h, double inc_cost, double tgt_cost) {
    WidthDecision d;
    d.finite_current_certificate = cert > 0.0 && tol > 0.0 && cert > tol && std::isfinite(cert) && std::isfinite(tol);
    d.history_mature = h.active_width == active && h.count >= 2;
    if (!d.finite_current_certificate || !d.history_mature || depth < 0 || max_depth < depth || rank < 0 ||
        limit < rank || active <= 0 || inc <= 0 || inc > kIntMax - active)
        return d;
    const long long recent_levels = static_cast<long long>(h.levels[0]) + h.levels[1];
    const ld recent = static_cast<ld>(h.lc[0]) + static_cast<ld>(h.lc[1]);
    if (recent_levels <= 0 || !(recent >= 0.0L) || !std::isfinite(recent)) return d;
    d.progress_stalled = recent == 0.0L;
    d.recent_level_lc = d.progress_stalled ? 0.0 : static_cast<double>(recent / static_cast<ld>(recent_levels));
    d.remaining_lc = std::log(cert) - std::log(tol);
    if ((!d.progress_stalled && (!(d.recent_level_lc > 0.0) || !std::isfinite(d.recent_level_lc))) ||
        !(d.remaining_lc > 0.0) || !std::isfinite(d.remaining_lc))
        return d;
    if (!d.progress_stalled) {
        const ld raw = static_cast<ld>(d.remaining_lc) / static_cast<ld>(d.recent_level_lc);
        if (!(raw > 0.0L) || !std::isfinite(raw)) return d;
        d.forecast_levels = raw >= static_cast<ld>(kIntMax) ? kIntMax
                                                           : std::max(1, static_cast<int>(std::ceil(raw)));
    }
    d.prior_history_mature = h.prior_levels > 0 && h.prior_lc > 0.0 && std::isfinite(h.prior_lc);
    if (d.prior_history_mature) {
        d.prior_level_lc = h.prior_lc / static_cast<double>(h.prior_levels);
        d.persistent_slowdown = std::isfinite(d.prior_level_lc) && h.level_lc[0] > 0.0 && h.level_lc[1] > 0.0 &&
                                h.level_lc[0] < d.prior_level_lc && h.level_lc[1] < d.prior_level_lc;
    }
    const int remaining_basis = limit - rank;
    const int target = active + inc;
    const int target_remaining = std::min(max_depth - depth, remaining_basis / target);
    const long long raw_h = d.progress_stalled ? static_cast<long long>(target) * target_remaining
                                               : static_cast<long long>(active) * d.forecast_levels;
    const int matched = static_cast<int>(std::min<long long>(raw_h, remaining_basis));
    WidthDecision c = horizon_candidate(inc, matched, depth, max_depth, rank, limit, active, h.op_s, h.loc_s, 2,
                                        inc_cost, tgt_cost, false);
    c.finite_current_certificate = d.finite_current_certificate;
    c.history_mature = d.history_mature;
    c.prior_history_mature = d.prior_history_mature;
    c.persistent_slowdown = d.persistent_slowdown;
    c.progress_stalled = d.progress_stalled;
    c.recent_level_lc = d.recent_level_lc;
    c.prior_level_lc = d.prior_level_lc;
    c.remaining_lc = d.remaining_lc;
    c.forecast_levels = d.forecast_levels;
    c.long_horizon_pressure = c.target_remaining_levels > 0 &&
                              (c.progress_stalled || d.forecast_levels > c.target_remaining_levels);
    const bool positive_recent = h.level_lc[0] > 0.0 && h.level_lc[1] > 0.0;
    c.recent_slowdown = positive_recent && h.level_lc[1] < 0.5 * h.level_lc[0];
    const bool established_stall = c.progress_stalled && (active > 1 || c.prior_history_mature);
    const bool scalar_needed = active == 1 && positive_recent &&
                               (c.recent_slowdown ||
                                (c.prior_history_mature && (c.persistent_slowdown || c.long_horizon_pressure)));
    const bool block_needed = active > 1 && (positive_recent || established_stall) &&
                              (c.persistent_slowdown || c.long_horizon_pressure || c.pass_reduction_admissible);
    c.widening_necessary = scalar_needed || block_needed;
    c.candidate = c.candidate && c.widening_necessary;
    return c;
}

struct MatchedHorizonPlan {
    int end_rank = -1, origin_width = -1, active_width = -1;
    double starting_certificate = kInf;
    double op_s[2] = {kInf, kInf};
    double loc_s[2] = {kInf, kInf};
    int samples = 0;
};

inline void reset(MatchedHorizonPlan& p) { p = MatchedHorizonPlan{}; }

inline int remaining(const MatchedHorizonPlan& p, int rank) {
    if (p.end_rank < 0 || rank < 0 || rank >= p.end_rank) return 0;
    return p.end_rank - rank;
}

inline bool extend_to_limit(MatchedHorizonPlan& p, int rank, int limit) {
    if (p.end_rank < 0 || rank < 0 || limit <= rank || rank < p.end_rank) return false;
    p.end_rank = limit;
    return true;
}

inline bool progress_stalled(const MatchedHorizonPlan& p, double cert) {
    return p.starting_certificate > 0.0 && cert > 0.0 && std::isfinite(p.starting_certificate) && std::isfinite(cert) &&
           std::log(p.starting_certificate) - std::log(cert) < std::log(2.0);
}

inline bool start_plan(MatchedHorizonPlan& p, int rank, int limit, int horizon, int origin, int target, double cert) {
    if (rank < 0 || limit < rank || horizon <= 0 || horizon > limit - rank || origin <= 0 || target <= 0 ||
        !(cert > 0.0) || !std::isfinite(cert)) {
        reset(p);
        return false;
    }
    p = MatchedHorizonPlan{};
    p.end_rank = rank + horizon;
    p.origin_width = origin;
    p.active_width = target;
    p.starting_certificate = cert;
    return true;
}

inline void reset_costs(MatchedHorizonPlan& p, int active) {
    p.active_width = active;
    p.op_s[0] = p.op_s[1] = kInf;
    p.loc_s[0] = p.loc_s[1] = kInf;
    p.samples = 0;
}

inline bool record_cost(MatchedHorizonPlan& p, int w0, int w1, int levels, double op_s, double loc_s) {
    if (p.end_rank < 0 || w0 <= 0 || w0 != w1 || w0 != p.active_width || levels <= 0 || op_s < 0.0 ||
        loc_s < 0.0 || !std::isfinite(op_s) || !std::isfinite(loc_s) || !(op_s + loc_s > 0.0) || p.samples >= 2)
        return false;
    const int i = p.samples++;
    p.op_s[i] = op_s / levels;
    p.loc_s[i] = loc_s / levels;
    return true;
}

// Certificate checkpoint cadence (certificate_checkpoint_schedule.hpp:15-131).
inline bool is_checkpoint(int it, double lambda) {
    if (it <= 1) return true;
    int odd = it;
    while (odd % 2 == 0) odd /= 2;
    return odd == 1 || (lambda == 0.0 && odd == 3);
}

inline int predicted_crossing(double prev, double cur, int prev_it, int cur_it, double tol) {
    if (!(prev > cur) || !(cur > tol) || !(tol > 0.0) || !std::isfinite(prev) || !std::isfinite(cur) ||
        !std::isfinite(tol))
        return kIntMax;
    const ld dist = std::log(static_cast<ld>(cur)) - std::log(static_cast<ld>(tol));
    const ld contraction = std::log(static_cast<ld>(prev)) - std::log(static_cast<ld>(cur));
    const ld off = std::ceil(static_cast<ld>(cur_it - prev_it) * dist / contraction);
    if (!(off > 0.0L) || !std::isfinite(off) || off > static_cast<ld>(kIntMax - cur_it)) return kIntMax;
    return cur_it + static_cast<int>(off);
}

class CheckpointSchedule {
public:
    CheckpointSchedule(double lambda, double tol) : lambda_(lambda), tol_(tol) {}
    bool should_evaluate(int it) const { return is_checkpoint(it, lambda_) || it >= predicted_; }
    template <typename R>
    void record(int it, const R& r) {
        predicted_ = kIntMax;
        if (is_checkpoint(it, lambda_) && has_prev_ && it > prev_it_) {
            if (lambda_ > 0.0) {
                predicted_ = predicted_crossing(prev_rel_, r.relative_energy_error_upper_bound, prev_it_, it, tol_);
            } else {
                predicted_ = std::min(
                    predicted_crossing(prev_comp_, r.compatible_backward_error_upper_bound, prev_it_, it, tol_),
                    predicted_crossing(prev_ls_, r.least_squares_backward_error_upper_bound, prev_it_, it, tol_));
            }
        }
        prev_it_ = it;
        prev_comp_ = r.compatible_backward_error_upper_bound;
        prev_ls_ = r.least_squares_backward_error_upper_bound;
        prev_rel_ = r.relative_energy_error_upper_bound;
        has_prev_ = true;
    }

private:
    double lambda_, tol_;
    int predicted_ = kIntMax, prev_it_ = 0;
    double prev_comp_ = 0.0, prev_ls_ = 0.0, prev_rel_ = 0.0;
    bool has_prev_ = false;
};

}  // namespace amflsx::ctl
