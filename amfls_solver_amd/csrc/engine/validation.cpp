#include "validation.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace amflsx {
namespace {

bool model_available(const ValidationModel& m) {
    return fp::environment_ok() && std::isfinite(m.upper) && m.upper >= 0.0 && std::isfinite(m.lower) &&
           m.lower >= 0.0 && m.lower <= m.upper && m.steps_apply > 0 && m.steps_transpose > 0;
}

double underflow_radius(long long ops, int64_t out) {
    if (ops <= 0 || out <= 0) return 0.0;
    return fp::up_ld(static_cast<long double>(ops) * std::sqrt(static_cast<long double>(out)) *
                     static_cast<long double>(fp::kDenormMin));
}

double callback_radius(const ValidationModel& m, bool transpose, double in_upper, int64_t out) {
    if (!model_available(m) || !std::isfinite(in_upper) || in_upper < 0.0) return kInfD;
    if (m.upper == 0.0 || in_upper == 0.0) return 0.0;
    const long long ops = transpose ? m.steps_transpose : m.steps_apply;
    const double main = fp::up_mul(fp::gamma64(ops), fp::up_mul(m.upper, in_upper));
    return fp::up_add(main, underflow_radius(ops, out));
}

double one_op_radius(double stored_upper, bool can_underflow, int64_t out) {
    if (!std::isfinite(stored_upper) || stored_upper < 0.0) return kInfD;
    double r = fp::up_mul(fp::gamma64(2), stored_upper);
    if (can_underflow) r = fp::up_add(r, underflow_radius(1, out));
    return r;
}

fp::Interval enlarge(fp::Interval s, double radius) {
    if (!std::isfinite(radius) || radius < 0.0) return {};
    return {fp::down_diff(s.lower, radius), fp::up_add(s.upper, radius)};
}

fp::Interval augmented(fp::Interval p, fp::Interval q, double lambda) {
    const double lo2 = fp::down_add(fp::down_mul(p.lower, p.lower),
                                    fp::down_mul(lambda, fp::down_mul(q.lower, q.lower)));
    const double hi2 = fp::up_add(fp::up_mul(p.upper, p.upper), fp::up_mul(lambda, fp::up_mul(q.upper, q.upper)));
    return {fp::down_sqrt(lo2), fp::up_sqrt(hi2)};
}

double certified_ratio(double num, double den) {
    if (!std::isfinite(num)) return kInfD;
    if (num == 0.0) return 0.0;
    if (!std::isfinite(den) || den <= 0.0) return kInfD;
    return num / den;
}

double relative_energy(double abs_bound, double energy) {
    if (!std::isfinite(energy)) return kInfD;
    if (abs_bound == 0.0) return 0.0;
    if (!std::isfinite(abs_bound) || energy <= 0.0) return kInfD;
    const double s = abs_bound / energy;
    if (!std::isfinite(s) || s >= 1.0) return kInfD;
    return s / (1.0 - s);
}

double ld_gamma(long long ops) {
    if (ops < 0) return kInfD;
    const long double p = static_cast<long double>(ops) * (0.5L * std::numeric_limits<long double>::epsilon());
    if (!(p < 1.0L)) return kInfD;
    return fp::up_ld(p / (1.0L - p));
}

double ld_underflow(long long ops, int64_t out) {
    if (ops <= 0 || out <= 0) return 0.0;
    return fp::up_ld(static_cast<long double>(ops) * std::sqrt(static_cast<long double>(out)) *
                     std::numeric_limits<long double>::denorm_min());
}

double ld_callback_radius(const ValidationModel& m, long long len, double in_upper, int64_t out) {
    if (!model_available(m) || len <= 0 || !std::isfinite(in_upper) || in_upper < 0.0) return kInfD;
    if (m.upper == 0.0 || in_upper == 0.0) return 0.0;
    const long long ops = 2 * len;
    return fp::up_add(fp::up_mul(ld_gamma(ops), fp::up_mul(m.upper, in_upper)), ld_underflow(ops, out));
}

double ld_one_op_radius(double stored_upper, bool can_underflow, int64_t out) {
    if (!std::isfinite(stored_upper) || stored_upper < 0.0) return kInfD;
    double r = fp::up_mul(ld_gamma(2), stored_upper);
    if (can_underflow) r = fp::up_add(r, ld_underflow(1, out));
    return r;
}

struct HostStats { double max = 0, sum = 0, bad = 0; };
HostStats host_stats(const std::vector<double>& v) {
    HostStats s;
    for (double x : v) {
        const double a = std::abs(x);
        if (!std::isfinite(a)) { s.bad += 1.0; continue; }
        s.max = std::max(s.max, a);
    }
    int e = 0;
    if (s.max > 0.0) (void)std::frexp(s.max, &e);
    for (double x : v) {
        const double a = std::abs(x);
        if (!std::isfinite(a)) continue;
        const double sc = std::ldexp(a, -e);
        s.sum = std::fma(sc, sc, s.sum);
    }
    return s;
}

}  // namespace

bool candidate_validation_failed_numerically(const LeastSquaresResult& r) {
    return r.status == SolverStatus::numerical_breakdown || r.stop_reason == StopReason::numerical_breakdown;
}

void evaluate_accuracy_contract(LeastSquaresResult& res, double bnorm, double lambda, double tol,
                                const double* ridge_bound) {
    const double root = std::sqrt(lambda);
    res.augmented_residual_norm = std::hypot(res.residual_norm, root * res.solution_norm);
    res.augmented_operator_norm_lower_bound = std::hypot(res.operator_norm_lower_bound, root);
    res.compatible_backward_error_upper_bound = certified_ratio(
        res.augmented_residual_norm, res.augmented_operator_norm_lower_bound * res.solution_norm + bnorm);
    res.least_squares_backward_error_upper_bound = certified_ratio(
        res.normal_residual_norm, res.augmented_operator_norm_lower_bound * res.augmented_residual_norm);
    res.backward_error_upper_bound =
        std::min(res.compatible_backward_error_upper_bound, res.least_squares_backward_error_upper_bound);
    const double nden = (res.operator_norm_lower_bound * res.operator_norm_lower_bound + lambda) * res.solution_norm +
                        res.operator_norm_lower_bound * bnorm;
    res.relative_normal_residual_upper_bound = certified_ratio(res.normal_residual_norm, nden);
    if (lambda > 0.0) {
        const double req = ridge_bound ? *ridge_bound : res.normal_residual_norm / root;
        res.energy_error_upper_bound = req >= 0.0 ? req : kInfD;
        res.objective_gap_upper_bound = res.energy_error_upper_bound * res.energy_error_upper_bound;
        res.relative_energy_error_upper_bound = relative_energy(res.energy_error_upper_bound, res.solution_energy_norm);
    } else if (res.normal_residual_norm == 0.0) {
        res.energy_error_upper_bound = res.relative_energy_error_upper_bound = res.objective_gap_upper_bound = 0.0;
    } else {
        res.energy_error_upper_bound = res.relative_energy_error_upper_bound = res.objective_gap_upper_bound = kInfD;
    }
    auto fin = [](double v) { return std::isfinite(v) && v >= 0.0; };
    if (!(fin(bnorm) && fin(res.residual_norm) && fin(res.augmented_residual_norm) &&
          fin(res.normal_residual_norm) && fin(res.solution_norm) && fin(res.solution_energy_norm) &&
          fin(res.objective) && fin(res.operator_norm_lower_bound) &&
          fin(res.augmented_operator_norm_lower_bound))) {
        res.status = SolverStatus::numerical_breakdown;
        res.stop_reason = StopReason::numerical_breakdown;
        return;
    }
    const bool passed = lambda > 0.0 ? res.relative_energy_error_upper_bound <= tol
                                     : res.backward_error_upper_bound <= tol;
    res.status = passed ? SolverStatus::success : SolverStatus::work_limit;
    if (res.normal_residual_norm == 0.0) res.stop_reason = StopReason::exact_stationarity;
    else if (passed && lambda > 0.0) res.stop_reason = StopReason::relative_energy_error;
    else if (passed && res.compatible_backward_error_upper_bound <= res.least_squares_backward_error_upper_bound)
        res.stop_reason = StopReason::compatible_backward_error;
    else if (passed) res.stop_reason = StopReason::least_squares_backward_error;
    else res.stop_reason = StopReason::maximum_depth;
}

void evaluate_certified_contract(LeastSquaresResult& res, const Envelope& env, double lambda, double tol,
                                 const double* ridge_bound) {
    if (!env.certified) {
        res.compatible_backward_error_upper_bound = res.least_squares_backward_error_upper_bound =
            res.backward_error_upper_bound = res.relative_normal_residual_upper_bound =
                res.energy_error_upper_bound = res.relative_energy_error_upper_bound =
                    res.objective_gap_upper_bound = kInfD;
        res.status = SolverStatus::work_limit;
        res.stop_reason = StopReason::maximum_depth;
        return;
    }
    const fp::Interval sol{env.solution_lower, env.solution_upper};
    const fp::Interval resid{env.residual_lower, env.residual_upper};
    const fp::Interval aug = augmented(resid, sol, lambda);
    const double lo = res.operator_norm_lower_bound;
    res.augmented_operator_norm_lower_bound = fp::down_sqrt(fp::down_add(fp::down_mul(lo, lo), lambda));
    const double cden = fp::down_add(fp::down_mul(res.augmented_operator_norm_lower_bound, env.solution_lower),
                                     env.right_hand_side_lower);
    res.compatible_backward_error_upper_bound = fp::up_div(aug.upper, cden);
    const double lden = fp::down_mul(res.augmented_operator_norm_lower_bound, aug.lower);
    res.least_squares_backward_error_upper_bound = fp::up_div(env.gradient_upper, lden);
    res.backward_error_upper_bound =
        std::min(res.compatible_backward_error_upper_bound, res.least_squares_backward_error_upper_bound);
    const double nden = fp::down_add(
        fp::down_mul(fp::down_add(fp::down_mul(lo, lo), lambda), env.solution_lower),
        fp::down_mul(lo, env.right_hand_side_lower));
    res.relative_normal_residual_upper_bound = fp::up_div(env.gradient_upper, nden);
    if (lambda > 0.0) {
        const double root_lo = fp::down_sqrt(lambda);
        const double base = fp::up_div(env.gradient_upper, root_lo);
        res.energy_error_upper_bound = ridge_bound ? *ridge_bound : base;
        res.objective_gap_upper_bound = fp::up_mul(res.energy_error_upper_bound, res.energy_error_upper_bound);
        const double rden = fp::down_diff(env.solution_energy_lower, res.energy_error_upper_bound);
        res.relative_energy_error_upper_bound = fp::up_div(res.energy_error_upper_bound, rden);
    } else if (env.gradient_upper == 0.0) {
        res.energy_error_upper_bound = res.relative_energy_error_upper_bound = res.objective_gap_upper_bound = 0.0;
    } else {
        res.energy_error_upper_bound = res.relative_energy_error_upper_bound = res.objective_gap_upper_bound = kInfD;
    }
    const bool passed = env.gradient_upper == 0.0 ||
                        (lambda > 0.0 ? res.relative_energy_error_upper_bound <= tol
                                      : res.backward_error_upper_bound <= tol);
    res.status = passed ? SolverStatus::success : SolverStatus::work_limit;
    if (env.gradient_upper == 0.0) res.stop_reason = StopReason::exact_stationarity;
    else if (passed && lambda > 0.0) res.stop_reason = StopReason::relative_energy_error;
    else if (passed && res.compatible_backward_error_upper_bound <= res.least_squares_backward_error_upper_bound)
        res.stop_reason = StopReason::compatible_backward_error;
    else if (passed) res.stop_reason = StopReason::least_squares_backward_error;
    else res.stop_reason = StopReason::maximum_depth;
}

IterationRecord make_iteration_record(const LeastSquaresResult& r, int epoch, int depth, int aux_width) {
    const RunStatistics& s = r.statistics;
    IterationRecord rec;
    rec.epoch = epoch;
    rec.depth = depth;
    rec.auxiliary_width = aux_width;
    rec.basis_rank = r.basis_rank;
    rec.a_columns = s.a_columns;
    rec.at_columns = s.at_columns;
    rec.a_block_calls = s.a_block_calls;
    rec.at_block_calls = s.at_block_calls;
    rec.contract_status = r.status;
    rec.contract_stop_reason = r.stop_reason;
    rec.contract_passed = r.status == SolverStatus::success;
    rec.ridge_correction_disposition = r.ridge_correction_disposition;
    rec.objective = r.objective;
    rec.residual_norm = r.residual_norm;
    rec.augmented_residual_norm = r.augmented_residual_norm;
    rec.normal_residual_norm = r.normal_residual_norm;
    rec.operator_norm_lower_bound = r.operator_norm_lower_bound;
    rec.compatible_backward_error_upper_bound = r.compatible_backward_error_upper_bound;
    rec.least_squares_backward_error_upper_bound = r.least_squares_backward_error_upper_bound;
    rec.backward_error_upper_bound = r.backward_error_upper_bound;
    rec.relative_energy_error_upper_bound = r.relative_energy_error_upper_bound;
    rec.relative_normal_residual_upper_bound = r.relative_normal_residual_upper_bound;
    rec.solution_norm = r.solution_norm;
    rec.solution_energy_norm = r.solution_energy_norm;
    rec.augmented_operator_norm_lower_bound = r.augmented_operator_norm_lower_bound;
    rec.energy_error_upper_bound = r.energy_error_upper_bound;
    rec.objective_gap_upper_bound = r.objective_gap_upper_bound;
    rec.ridge_base_energy_error_upper_bound = r.ridge_base_energy_error_upper_bound;
    rec.ridge_corrected_energy_error_upper_bound = r.ridge_corrected_energy_error_upper_bound;
    rec.ridge_correction_gamma = r.ridge_correction_gamma;
    rec.ridge_correction_z_h_z = r.ridge_correction_z_h_z;
    rec.ridge_correction_two_abs_z_t_q = r.ridge_correction_two_abs_z_t_q;
    rec.ridge_correction_q_norm_squared = r.ridge_correction_q_norm_squared;
    return rec;
}

ValidationEngine::ValidationEngine(const DeviceOperator& op, const double* b_dev, Context& ctx,
                                   RunStatistics& stats)
    : op_(op), b_dev_(b_dev), ctx_(ctx), stats_(stats), m_(op.rows()), n_(op.cols()) {
    ax_.reset(m_);
    r_.reset(m_);
    atr_.reset(n_);
    g_.reset(n_);
    d_.reset(n_);
    ad_.reset(m_);
    td_.reset(n_);
    hd_.reset(n_);
    sb_ = stats_of(b_dev_, m_);
    bnorm_ = nominal(sb_);
    b_iv_ = interval(sb_, m_);
}

ValidationEngine::Stats3 ValidationEngine::stats_of(const double* v, int64_t n) {
    double out[3];
    ctx_.block_stats<double>(n, 1, v, 1, out);
    return {out[0], out[1], out[2]};
}

void ValidationEngine::stats_of_many(const double* const* v, const int64_t* n, int count, Stats3* out) {
    // device slots 16.. of the context's statistics buffer (slot 8 holds a
    // dot product of the ridge correction)
    constexpr int kFirst = 16;
    if (kFirst + 3 * count > static_cast<int>(ctx_.stats.size()))
        throw std::logic_error("too many statistics in one batch");
    for (int i = 0; i < count; ++i) ctx_.block_stats_async<double>(n[i], 1, v[i], 1, ctx_.stats.get() + kFirst + 3 * i);
    double host[48];
    ctx_.download(host, ctx_.stats.get() + kFirst, static_cast<size_t>(3 * count));
    for (int i = 0; i < count; ++i) out[i] = {host[3 * i], host[3 * i + 1], host[3 * i + 2]};
}

void ValidationEngine::count(bool transpose, bool base, double sec) {
    RunStatistics& s = stats_;
    if (transpose) {
        s.at_seconds += sec;
        s.at_columns += 1;
        s.validation_at_columns += 1;
        s.at_block_calls += 1;
        s.validation_at_block_calls += 1;
        if (base) { s.base_validation_at_columns += 1; s.base_validation_at_block_calls += 1; s.base_validation_seconds += sec; }
        else { s.ridge_correction_at_columns += 1; s.ridge_correction_at_block_calls += 1; s.ridge_correction_validation_seconds += sec; }
    } else {
        s.a_seconds += sec;
        s.a_columns += 1;
        s.validation_a_columns += 1;
        s.a_block_calls += 1;
        s.validation_a_block_calls += 1;
        if (base) { s.base_validation_a_columns += 1; s.base_validation_a_block_calls += 1; s.base_validation_seconds += sec; }
        else { s.ridge_correction_a_columns += 1; s.ridge_correction_a_block_calls += 1; s.ridge_correction_validation_seconds += sec; }
    }
    s.validation_seconds = s.base_validation_seconds + s.ridge_correction_validation_seconds;
}

void ValidationEngine::apply_counted(const double* in, double* out, bool transpose, bool base) {
    Clock& clk = ctx_.clock;
    if (!clk.synthetic() && op_.times_own_kernels()) {
        // the product's kernels stamp their own time, added to the operator
        // and validation seconds when the timer is flushed: no waits here
        RunStatistics& st = stats_;
        const StampScope stamp_scope{op_};
        op_.stamp_next(ctx_.timer.kernel_stamp(ctx_.stream, transpose ? &st.at_seconds : &st.a_seconds,
                                               base ? &st.base_validation_seconds
                                                    : &st.ridge_correction_validation_seconds));
        if (transpose) op_.apply_t(ctx_.stream, in, 1, 1, out, 1);
        else op_.apply(ctx_.stream, in, 1, 1, out, 1);
        count(transpose, base, 0.0);
        return;
    }
    if (!clk.synthetic()) ctx_.sync();
    const double start = clk.now();
    clk.charge_operator(op_.relative_block_product_cost(1));
    if (transpose) op_.apply_t(ctx_.stream, in, 1, 1, out, 1);
    else op_.apply(ctx_.stream, in, 1, 1, out, 1);
    if (!clk.synthetic()) ctx_.sync();
    count(transpose, base, clk.now() - start);
}

Envelope ValidationEngine::base_envelope(double lambda, const Stats3& sb, const Stats3& sx, const Stats3& sax,
                                         const Stats3& sr, const Stats3& satr, const Stats3& sg) {
    Envelope env;
    env.model = op_.validation_model();
    env.operator_norm_upper_bound = env.model.upper;
    if (!model_available(env.model)) return env;
    const fp::Interval bn = interval(sb, m_), xn = interval(sx, n_), axn = interval(sax, m_);
    const fp::Interval rn = interval(sr, m_), tn = interval(satr, n_), gn = interval(sg, n_);
    const double apply_err = callback_radius(env.model, false, xn.upper, m_);
    const double sub_err = one_op_radius(rn.upper, axn.upper > 0.0 || bn.upper > 0.0, m_);
    const double res_err = fp::up_add(apply_err, sub_err);
    const fp::Interval true_res = enlarge(rn, res_err);
    const double t_err = callback_radius(env.model, true, rn.upper, n_);
    const double ridge_err = one_op_radius(gn.upper, lambda > 0.0 && xn.upper > 0.0, n_);
    const double prop = fp::up_mul(env.operator_norm_upper_bound, res_err);
    const double g_err = fp::up_add(fp::up_add(t_err, prop), ridge_err);
    const fp::Interval true_g = enlarge(gn, g_err);
    const fp::Interval true_ax = enlarge(axn, apply_err);
    double op_lo = env.model.lower;
    if (xn.upper > 0.0) op_lo = std::max(op_lo, fp::down_div(fp::down_diff(axn.lower, apply_err), xn.upper));
    if (rn.upper > 0.0) op_lo = std::max(op_lo, fp::down_div(fp::down_diff(tn.lower, t_err), rn.upper));
    const fp::Interval energy = augmented(true_ax, xn, lambda);
    env.residual_lower = true_res.lower;
    env.residual_upper = true_res.upper;
    env.gradient_upper = true_g.upper;
    env.solution_lower = xn.lower;
    env.solution_upper = xn.upper;
    env.right_hand_side_lower = std::max(0.0, bn.lower);
    env.applied_solution_lower = std::max(0.0, true_ax.lower);
    env.applied_solution_upper = true_ax.upper;
    env.solution_energy_lower = energy.lower;
    env.gradient_error_radius = g_err;
    env.certified = std::isfinite(env.residual_upper) && std::isfinite(env.gradient_upper) &&
                    std::isfinite(env.solution_upper) && std::isfinite(env.operator_norm_upper_bound);
    env.operator_norm_lower_bound = op_lo;
    return env;
}

bool ValidationEngine::refine(const double* x_dev, double lambda, Envelope& env, Stats3& sax, Stats3& sr,
                              Stats3& sg, double& ax_nominal) {
    // reference candidate_validation.cpp:951-1149 (long double on the host)
    if (!op_.supports_refinement()) return false;
    const ValidationModel model = op_.validation_model();
    if (!model_available(model)) return false;
    std::vector<double> hx(n_), hb(m_);
    ctx_.download(hx.data(), x_dev, n_);
    ctx_.download(hb.data(), b_dev_, m_);
    const HostStats hsb = host_stats(hb), hsx = host_stats(hx);
    const fp::Interval bn = fp::norm_interval(hsb.max, hsb.sum, hsb.bad, m_);
    const fp::Interval xn = fp::norm_interval(hsx.max, hsx.sum, hsx.bad, n_);
    if (!std::isfinite(bn.upper) || !std::isfinite(xn.upper)) return false;
    std::vector<long double> lx(hx.begin(), hx.end()), lax(m_), lr(m_), lt(n_), lg(n_);
    Clock& clk = ctx_.clock;
    double start = clk.now();
    clk.charge_operator(op_.relative_block_product_cost(1));
    const bool ok_a = op_.refine_apply(lx.data(), lax.data());
    count(false, true, clk.now() - start);
    if (!ok_a) return false;
    std::vector<double> ax(m_), r(m_), atr(n_), g(n_);
    for (int64_t i = 0; i < m_; ++i) {
        if (!std::isfinite(lax[i])) return false;
        ax[i] = static_cast<double>(lax[i]);
        if (!std::isfinite(ax[i])) return false;
    }
    const HostStats hax = host_stats(ax);
    const fp::Interval axn = fp::norm_interval(hax.max, hax.sum, hax.bad, m_);
    const double ax_cast = one_op_radius(axn.upper, true, m_);
    const double ax_acc = ld_callback_radius(model, n_, xn.upper, m_);
    const double ax_err = fp::up_add(ax_acc, ax_cast);
    for (int64_t i = 0; i < m_; ++i) {
        lr[i] = lax[i] - static_cast<long double>(hb[i]);
        r[i] = static_cast<double>(lr[i]);
        if (!std::isfinite(lr[i]) || !std::isfinite(r[i])) return false;
    }
    const HostStats hr = host_stats(r);
    const fp::Interval rn = fp::norm_interval(hr.max, hr.sum, hr.bad, m_);
    const double r_cast = one_op_radius(rn.upper, true, m_);
    const double lr_upper = fp::up_add(rn.upper, r_cast);
    const double r_sub = ld_one_op_radius(lr_upper, axn.upper > 0.0 || bn.upper > 0.0, m_);
    const double lr_err = fp::up_add(ax_acc, r_sub);
    const double r_err = fp::up_add(lr_err, r_cast);
    start = clk.now();
    clk.charge_operator(op_.relative_block_product_cost(1));
    const bool ok_t = op_.refine_apply_t(lr.data(), lt.data());
    count(true, true, clk.now() - start);
    if (!ok_t) return false;
    for (int64_t j = 0; j < n_; ++j) {
        if (!std::isfinite(lt[j])) return false;
        atr[j] = static_cast<double>(lt[j]);
        if (!std::isfinite(atr[j])) return false;
    }
    const HostStats ht = host_stats(atr);
    const fp::Interval tn = fp::norm_interval(ht.max, ht.sum, ht.bad, n_);
    const double t_cast = one_op_radius(tn.upper, true, n_);
    const double t_acc = ld_callback_radius(model, m_, lr_upper, n_);
    for (int64_t j = 0; j < n_; ++j) {
        lg[j] = std::fma(static_cast<long double>(lambda), lx[j], lt[j]);
        g[j] = static_cast<double>(lg[j]);
        if (!std::isfinite(lg[j]) || !std::isfinite(g[j])) return false;
    }
    const HostStats hg = host_stats(g);
    const fp::Interval gn = fp::norm_interval(hg.max, hg.sum, hg.bad, n_);
    const double g_cast = one_op_radius(gn.upper, true, n_);
    const double lg_upper = fp::up_add(gn.upper, g_cast);
    const double ridge_err = ld_one_op_radius(lg_upper, lambda > 0.0 && xn.upper > 0.0, n_);
    const double prop = fp::up_mul(model.upper, lr_err);
    const double g_err = fp::up_add(fp::up_add(t_acc, prop), fp::up_add(ridge_err, g_cast));
    const fp::Interval true_ax = enlarge(axn, ax_err), true_r = enlarge(rn, r_err), true_g = enlarge(gn, g_err);
    double op_lo = model.lower;
    if (xn.upper > 0.0) op_lo = std::max(op_lo, fp::down_div(fp::down_diff(axn.lower, ax_err), xn.upper));
    if (rn.upper > 0.0) {
        const double conv = fp::up_mul(model.upper, r_cast);
        const double te = fp::up_add(fp::up_add(t_acc, t_cast), conv);
        op_lo = std::max(op_lo, fp::down_div(fp::down_diff(tn.lower, te), rn.upper));
    }
    const fp::Interval energy = augmented(true_ax, xn, lambda);
    Envelope e;
    e.model = model;
    e.operator_norm_upper_bound = model.upper;
    e.operator_norm_lower_bound = op_lo;
    e.residual_lower = true_r.lower;
    e.residual_upper = true_r.upper;
    e.gradient_upper = true_g.upper;
    e.solution_lower = xn.lower;
    e.solution_upper = xn.upper;
    e.right_hand_side_lower = std::max(0.0, bn.lower);
    e.applied_solution_lower = std::max(0.0, true_ax.lower);
    e.applied_solution_upper = true_ax.upper;
    e.solution_energy_lower = energy.lower;
    e.gradient_error_radius = g_err;
    e.certified = std::isfinite(e.residual_upper) && std::isfinite(e.gradient_upper) &&
                  std::isfinite(e.solution_upper) && std::isfinite(e.operator_norm_upper_bound);
    if (!e.certified) return false;
    env = e;
    sax = {hax.max, hax.sum, hax.bad};
    sr = {hr.max, hr.sum, hr.bad};
    sg = {hg.max, hg.sum, hg.bad};
    ax_nominal = fp::nominal_norm(hax.max, hax.sum, hax.bad);
    // the refined gradient feeds the ridge correction
    AMFLS_HIP_CHECK(hipMemcpyAsync(g_.get(), g.data(), n_ * sizeof(double), hipMemcpyHostToDevice, ctx_.stream));
    AMFLS_HIP_CHECK(hipMemcpyAsync(ax_.get(), ax.data(), m_ * sizeof(double), hipMemcpyHostToDevice, ctx_.stream));
    AMFLS_HIP_CHECK(hipMemcpyAsync(r_.get(), r.data(), m_ * sizeof(double), hipMemcpyHostToDevice, ctx_.stream));
    ctx_.sync();
    return true;
}

LeastSquaresResult ValidationEngine::validate(const double* x_dev, double lambda, double tol) {
    LeastSquaresResult res;
    res.rows = static_cast<int>(m_);
    res.cols = static_cast<int>(n_);
    res.operator_norm_lower_bound = 0.0;
    apply_counted(x_dev, ax_.get(), false, true);
    sub_vec(ctx_.stream, m_, ax_.get(), b_dev_, r_.get());
    apply_counted(r_.get(), atr_.get(), true, true);
    fma_vec(ctx_.stream, n_, lambda, x_dev, atr_.get(), g_.get());
    Stats3 s5[5];
    {
        const double* const vs[5] = {x_dev, ax_.get(), r_.get(), atr_.get(), g_.get()};
        const int64_t ns[5] = {n_, m_, m_, n_, n_};
        stats_of_many(vs, ns, 5, s5);
    }
    Stats3 sx = s5[0], sax = s5[1], sr = s5[2], satr = s5[3], sg = s5[4];
    auto populate = [&](double axn) {
        res.residual_norm = nominal(sr);
        res.normal_residual_norm = nominal(sg);
        res.solution_norm = nominal(sx);
        res.solution_energy_norm = std::hypot(axn, std::sqrt(lambda) * res.solution_norm);
        res.objective = res.residual_norm * res.residual_norm + lambda * res.solution_norm * res.solution_norm;
    };
    populate(nominal(sax));
    evaluate_accuracy_contract(res, bnorm_, lambda, tol, nullptr);
    const bool finite = sx.bad == 0 && sax.bad == 0 && sr.bad == 0 && satr.bad == 0 && sg.bad == 0;
    if (!finite || candidate_validation_failed_numerically(res)) {
        res.status = SolverStatus::numerical_breakdown;
        res.stop_reason = StopReason::numerical_breakdown;
    } else {
        Envelope env = base_envelope(lambda, sb_, sx, sax, sr, satr, sg);
        res.operator_norm_lower_bound = env.operator_norm_lower_bound;
        evaluate_certified_contract(res, env, lambda, tol, nullptr);
        if (res.status != SolverStatus::success && env.certified &&
            env.gradient_error_radius >= res.normal_residual_norm) {
            double axn = 0.0;
            if (refine(x_dev, lambda, env, sax, sr, sg, axn)) {
                populate(axn);
                evaluate_accuracy_contract(res, bnorm_, lambda, tol, nullptr);
                res.operator_norm_lower_bound = env.operator_norm_lower_bound;
                evaluate_certified_contract(res, env, lambda, tol, nullptr);
            }
        }
        if (lambda > 0.0) {
            res.ridge_base_energy_error_upper_bound = res.energy_error_upper_bound;
            if (res.status == SolverStatus::success) {
                res.ridge_correction_disposition = RidgeCorrection::skipped_base_pass;
            } else if (!env.certified) {
                res.ridge_correction_disposition = RidgeCorrection::skipped_nonfinite_candidate;
            } else if (env.solution_energy_lower == 0.0) {
                res.ridge_correction_disposition = RidgeCorrection::skipped_zero_energy;
            } else if (!(env.gradient_upper > 0.0) || !std::isfinite(env.gradient_upper)) {
                res.ridge_correction_disposition = RidgeCorrection::skipped_nonfinite_candidate;
            } else {
                // one_step_ridge_energy_bound (candidate_validation.cpp:395-566)
                RidgeCorrection disp = RidgeCorrection::attempted_nonfinite;
                double bound = kNaN, gamma = kNaN, zhz = kNaN, cross2 = kNaN, qsq = kNaN;
                const double gnorm = res.normal_residual_norm;
                const ValidationModel& model = env.model;
                [&] {
                    if (!(gnorm > 0.0) || !std::isfinite(gnorm)) {
                        disp = RidgeCorrection::skipped_nonfinite_candidate;
                        return;
                    }
                    scale_vec(ctx_.stream, n_, 1.0 / gnorm, g_.get(), d_.get());
                    // reference divides elementwise (g / ||g||); one rounding either way.
                    // d = g / ||g|| is finite whenever g is (checked by the base
                    // pass), so A d is formed before d's norm is read: one wait for
                    // both statistics.
                    apply_counted(d_.get(), ad_.get(), false, false);
                    Stats3 s2[2];
                    {
                        const double* const vs[2] = {d_.get(), ad_.get()};
                        const int64_t ns[2] = {n_, m_};
                        stats_of_many(vs, ns, 2, s2);
                    }
                    const Stats3 sd = s2[0], sad = s2[1];
                    if (sd.bad > 0) { disp = RidgeCorrection::skipped_nonfinite_candidate; return; }
                    const fp::Interval dn = interval(sd, n_);
                    if (sad.bad > 0) { disp = RidgeCorrection::attempted_partial_failure; return; }
                    const fp::Interval adn = interval(sad, m_);
                    const double ad_err = callback_radius(model, false, dn.upper, m_);
                    apply_counted(ad_.get(), td_.get(), true, false);
                    const double t_err = callback_radius(model, true, adn.upper, n_);
                    fma_vec(ctx_.stream, n_, lambda, d_.get(), td_.get(), hd_.get());
                    // curvature = d . hd (double dot on device), read with hd's norm
                    gemm_tn<double>(ctx_.stream, n_, 1, 1, Strided{d_.get(), ValueType::f64, 1, 1}, hd_.get(), 1,
                                    ctx_.stats.get() + 8, ValueType::f64, 1, 1.0, ctx_.scratch);
                    ctx_.block_stats_async<double>(n_, 1, hd_.get(), 1, ctx_.stats.get() + 16);
                    double h11[11];  // statistics buffer slots 8 .. 18
                    ctx_.download(h11, ctx_.stats.get() + 8, 11);
                    const double curvature = h11[0];
                    const Stats3 shd{h11[8], h11[9], h11[10]};
                    if (shd.bad > 0) return;
                    const fp::Interval hdn = interval(shd, n_);
                    const double prop = fp::up_mul(env.operator_norm_upper_bound, ad_err);
                    const double hsum = one_op_radius(hdn.upper, lambda > 0.0 && dn.upper > 0.0, n_);
                    const double h_err = fp::up_add(fp::up_add(t_err, prop), hsum);
                    if (!std::isfinite(curvature)) return;
                    if (!(curvature > 0.0)) { disp = RidgeCorrection::attempted_invalid_curvature; return; }
                    const double step = gnorm / curvature;
                    if (!(step > 0.0) || !std::isfinite(step)) return;
                    fma_vec(ctx_.stream, n_, -step, hd_.get(), g_.get(), hd_.get());
                    const Stats3 sq = stats_of(hd_.get(), n_);
                    if (sq.bad > 0) return;
                    const fp::Interval qn = interval(sq, n_);
                    const double axpy_err = one_op_radius(qn.upper, gnorm > 0.0 && hdn.upper > 0.0, n_);
                    const double q_err = fp::up_add(fp::up_add(env.gradient_error_radius, fp::up_mul(step, h_err)),
                                                    axpy_err);
                    const double q_upper = fp::up_add(qn.upper, q_err);
                    const double ad_true = fp::up_add(adn.upper, ad_err);
                    const double step2 = fp::up_mul(step, step);
                    const double de2 = fp::up_add(fp::up_mul(ad_true, ad_true),
                                                  fp::up_mul(lambda, fp::up_mul(dn.upper, dn.upper)));
                    zhz = fp::up_mul(step2, de2);
                    const double cross = fp::up_mul(step, fp::up_mul(dn.upper, q_upper));
                    cross2 = fp::up_mul(2.0, cross);
                    qsq = fp::up_mul(q_upper, q_upper);
                    const double qe = fp::up_div(qsq, lambda);
                    const double corrected = fp::up_sqrt(fp::up_add(zhz, fp::up_add(cross2, qe)));
                    if (std::isnan(corrected)) { zhz = cross2 = qsq = kNaN; return; }
                    bound = corrected;
                    gamma = 1.0 / curvature;
                    disp = RidgeCorrection::attempted_not_improved;
                }();
                res.ridge_correction_disposition = disp;
                res.ridge_corrected_energy_error_upper_bound = bound;
                res.ridge_correction_gamma = gamma;
                res.ridge_correction_z_h_z = zhz;
                res.ridge_correction_two_abs_z_t_q = cross2;
                res.ridge_correction_q_norm_squared = qsq;
                if (disp == RidgeCorrection::attempted_not_improved && std::isfinite(bound) &&
                    bound < res.energy_error_upper_bound) {
                    res.ridge_correction_disposition = RidgeCorrection::attempted_improved;
                    evaluate_certified_contract(res, env, lambda, tol, &bound);
                }
            }
        }
    }
    if (lambda > 0.0 && res.status == SolverStatus::numerical_breakdown) {
        res.ridge_base_energy_error_upper_bound = std::isnan(res.normal_residual_norm) ? kNaN : kInfD;
        res.energy_error_upper_bound = res.ridge_base_energy_error_upper_bound;
        res.relative_energy_error_upper_bound = kInfD;
        res.objective_gap_upper_bound = kInfD;
        res.ridge_correction_disposition = RidgeCorrection::skipped_nonfinite_candidate;
    }
    res.solution.resize(n_);
    ctx_.download(res.solution.data(), x_dev, n_);
    // the stream is idle after the download: resolve the product stamps so
    // the statistics carry this validation's operator seconds
    ctx_.timer.flush();
    stats_.validation_seconds = stats_.base_validation_seconds + stats_.ridge_correction_validation_seconds;
    res.statistics = stats_;
    return res;
}

}  // namespace amflsx
