#include "refine.hpp"

#include <algorithm>
#include <cmath>
#include <memory>

#include "fp_bounds.hpp"
#include "validation.hpp"

namespace amflsx {

namespace {

double device_norm(Context& ctx, const double* p, int64_t len) {
    double st[3];
    ctx.block_stats<double>(len, 1, p, 1, st);
    return fp::nominal_norm(st[0], st[1], st[2]);
}

double host_sq(const std::vector<double>& a) {
    long double s = 0.0L;
    for (double v : a) s += static_cast<long double>(v) * v;
    return static_cast<double>(s);
}

void host_scale(std::vector<double>& a, double f) {
    for (double& v : a) v *= f;
}

}  // namespace

BasisPreconditioner::BasisPreconditioner(Context& ctx, const BasisCapture& basis, double lambda)
    : ctx_(ctx), basis_(basis), n_(basis.n), k_(basis.k), complement_(basis.k < basis.n) {
    // Diagonals of R at FP32 noise level are kept (they are ~u32 ||A||);
    // only exact zeros are raised, so R^{-1} exists.
    raised_ = staircase_r(basis.T, lambda, 1e-10, R_);
    sigma_min_ = band_min_singular(R_, 8);
    if (complement_) {
        // The unexplored directions are scaled like the smallest explored
        // one (for lambda > 0, sigma_min(R) >= sqrt(lambda)).
        const double floor = std::sqrt(lambda);
        const double rho = std::max(sigma_min_, floor);
        tau_ = rho > 0.0 && std::isfinite(rho) ? 1.0 / rho : 1.0;
    }
    proj_.reset(static_cast<size_t>(std::max(k_, 1)));
}

void BasisPreconditioner::apply(const std::vector<double>& a, const double* c, double* y) {
    hipStream_t s = ctx_.stream;
    const float* V = basis_.V.get();
    t_ = a;
    band_solve(R_, t_);
    const double* td = ctx_.upload(t_.data(), t_.size());
    basis_apply<float>(s, n_, k_, V, k_, td, y, 1.0, 0.0);
    if (complement_ && c) {
        basis_project<float>(s, n_, k_, V, k_, c, proj_.get(), ctx_.scratch);
        axpy_block<double>(s, n_, 1, tau_, c, 1, y, 1);
        basis_apply<float>(s, n_, k_, V, k_, proj_.get(), y, -tau_, 1.0);
    }
}

void BasisPreconditioner::apply_t(const double* g, std::vector<double>& a, double* c) {
    hipStream_t s = ctx_.stream;
    const float* V = basis_.V.get();
    basis_project<float>(s, n_, k_, V, k_, g, proj_.get(), ctx_.scratch);
    if (complement_ && c) {
        copy_block<double, double>(s, n_, 1, g, 1, c, 1, tau_);
        basis_apply<float>(s, n_, k_, V, k_, proj_.get(), c, -tau_, 1.0);
    }
    a.resize(static_cast<size_t>(k_));
    ctx_.download(a.data(), proj_.get(), static_cast<size_t>(k_));
    band_solve_t(R_, a);
}

LeastSquaresResult refine_preconditioned(const DeviceOperator& op, const double* b_dev, const SolveOptions& o,
                                         Context& ctx, RunStatistics& stats, const BasisCapture& basis,
                                         double* x_dev, std::vector<IterationRecord>& trace, RefineReport& report) {
    const double lambda = o.regularization, root = std::sqrt(lambda);
    const int64_t m = op.rows(), n = op.cols();
    hipStream_t s = ctx.stream;
    std::unique_ptr<AugmentedOperator> aug;
    if (lambda > 0.0) aug = std::make_unique<AugmentedOperator>(op, root, ctx);
    const DeviceOperator& M = aug ? static_cast<const DeviceOperator&>(*aug) : op;
    const int64_t mi = M.rows();
    ValidationEngine validator(op, b_dev, ctx, stats);
    BasisPreconditioner P(ctx, basis, lambda);
    const bool comp = P.has_complement();
    const int k = P.rank();
    // A basis that spans R^n makes the preconditioned system orthonormal to
    // about kappa u32 and a cycle of a few iterations gains several digits; a
    // partial basis leaves the complement to LSQR itself, which then runs
    // like FP64 LSQR on those directions (longer cycles).
    const int max_it = o.maximum_refinement_iterations > 0 ? o.maximum_refinement_iterations : 2000;
    DevBuf<double> u(static_cast<size_t>(mi)), u2(static_cast<size_t>(mi)), y(static_cast<size_t>(n)),
        g(static_cast<size_t>(n)), fe(static_cast<size_t>(n));
    DevBuf<double> vc, wc, zc, tc;
    if (comp) {
        vc.reset(static_cast<size_t>(n));
        wc.reset(static_cast<size_t>(n));
        zc.reset(static_cast<size_t>(n));
        tc.reset(static_cast<size_t>(n));
    }
    std::vector<double> va, wa, za, ta;

    // FP64 products of [A; sqrt(lambda) I] (events time them on the stream).
    auto product = [&](const double* in, double* out, bool transpose) {
        ctx.timer.begin(s);
        if (transpose) M.apply_t(s, in, 1, 1, out, 1);
        else M.apply(s, in, 1, 1, out, 1);
        ctx.timer.end(s, transpose ? &stats.at_seconds : &stats.a_seconds);
        if (transpose) {
            ++stats.at_columns; ++stats.at_block_calls; ++stats.iterative_at_columns; ++stats.iterative_at_block_calls;
        } else {
            ++stats.a_columns; ++stats.a_block_calls; ++stats.iterative_a_columns; ++stats.iterative_a_block_calls;
        }
    };
    // v = P^T M^T u - beta v (beta = 0: v = P^T M^T u); returns ||v||.
    auto transpose_step = [&](double beta) {
        product(u.get(), g.get(), true);
        P.apply_t(g.get(), ta, comp ? tc.get() : nullptr);
        if (beta != 0.0) {
            for (int i = 0; i < k; ++i) ta[i] -= beta * va[i];
            if (comp) axpy_block<double>(s, n, 1, -beta, vc.get(), 1, tc.get(), 1);
        }
        va.swap(ta);
        if (comp) std::swap(vc, tc);
        double sq = host_sq(va);
        if (comp) {
            const double nc = device_norm(ctx, vc.get(), n);
            sq += nc * nc;
        }
        return std::sqrt(sq);
    };
    // One LSQR cycle of at most `len` iterations on min ||M P z - u|| (u
    // holds the FP64 residual [b - A x; -sqrt(lambda) x] of the current x),
    // then x += P z.  Returns the iterations taken.
    auto cycle = [&](int len) {
        za.assign(static_cast<size_t>(k), 0.0);
        if (comp) fill<double>(s, zc.get(), n, 0.0);
        double beta = device_norm(ctx, u.get(), mi);
        if (!(beta > 0.0) || !std::isfinite(beta)) return 0;
        scale_vec(s, mi, 1.0 / beta, u.get(), u.get());
        double alpha = transpose_step(0.0);
        if (!(alpha > 0.0) || !std::isfinite(alpha)) return 0;
        host_scale(va, 1.0 / alpha);
        if (comp) scale_vec(s, n, 1.0 / alpha, vc.get(), vc.get());
        wa = va;
        if (comp) copy_block<double, double>(s, n, 1, vc.get(), 1, wc.get(), 1, 1.0);
        double rhobar = alpha, phibar = beta;
        int it = 0;
        while (it < len) {
            ++it;
            // beta u = M P v - alpha u
            P.apply(va, comp ? vc.get() : nullptr, y.get());
            product(y.get(), u2.get(), false);
            axpy_block<double>(s, mi, 1, -alpha, u.get(), 1, u2.get(), 1);
            std::swap(u, u2);
            beta = device_norm(ctx, u.get(), mi);
            const bool beta_ok = beta > 0.0 && std::isfinite(beta);
            double next_alpha = 0.0;
            if (beta_ok) {
                scale_vec(s, mi, 1.0 / beta, u.get(), u.get());
                // alpha v = P^T M^T u - beta v
                next_alpha = transpose_step(beta);
                if (next_alpha > 0.0 && std::isfinite(next_alpha)) {
                    host_scale(va, 1.0 / next_alpha);
                    if (comp) scale_vec(s, n, 1.0 / next_alpha, vc.get(), vc.get());
                }
            }
            const double rho = std::hypot(rhobar, beta_ok ? beta : 0.0);
            if (!(rho > 0.0) || !std::isfinite(rho)) break;
            const double c = rhobar / rho, sn = (beta_ok ? beta : 0.0) / rho;
            const double theta = sn * next_alpha;
            rhobar = -c * next_alpha;
            const double phi = c * phibar;
            phibar = sn * phibar;
            if (!std::isfinite(phi)) break;
            // z += (phi / rho) w;  w = v - (theta / rho) w
            const double zf = phi / rho, wf = -theta / rho;
            for (int i = 0; i < k; ++i) {
                za[i] += zf * wa[i];
                wa[i] = va[i] + wf * wa[i];
            }
            if (comp) {
                axpy_block<double>(s, n, 1, zf, wc.get(), 1, zc.get(), 1);
                copy_block<double, double>(s, n, 1, wc.get(), 1, wc.get(), 1, wf);
                axpy_block<double>(s, n, 1, 1.0, vc.get(), 1, wc.get(), 1);
            }
            if (!beta_ok || !(next_alpha > 0.0) || !std::isfinite(next_alpha)) break;
            alpha = next_alpha;
        }
        P.apply(za, comp ? zc.get() : nullptr, y.get());
        add_vec(s, n, y.get(), x_dev);
        return it;
    };

    LeastSquaresResult best;
    bool have = false;
    DevBuf<double> xbest(static_cast<size_t>(n));
    auto cert_of = [&](const LeastSquaresResult& r) {
        return lambda > 0.0 ? r.relative_energy_error_upper_bound : r.backward_error_upper_bound;
    };
    // certified first, then the smaller forward-error estimate, then the
    // smaller certificate
    auto better = [&](const LeastSquaresResult& r) {
        if (!have) return true;
        const bool rs = r.status == SolverStatus::success, bs = best.status == SolverStatus::success;
        if (rs != bs) return rs;
        const double re = r.forward_error_estimate, be = best.forward_error_estimate;
        if (std::isfinite(re) != std::isfinite(be)) return std::isfinite(re);
        if (std::isfinite(re) && re != be) return re < be;
        return cert_of(r) < cert_of(best);
    };
    // Certify x (FP64 validator), estimate its forward error ||P P^T g|| /
    // ||x|| (g = A^T (A x - b) + lambda x), and leave the FP64 residual
    // [b - A x; -sqrt(lambda) x] in u for the next cycle.
    auto certify = [&](int it, bool* reached) {
        LeastSquaresResult r = validator.validate(x_dev, lambda, o.tolerance);
        r.iterations = it;
        ++report.checkpoints;
        double est = kInfD;
        if (std::isfinite(r.solution_norm) && r.solution_norm > 0.0) {
            std::vector<double> ga;
            P.apply_t(validator.last_gradient(), ga, comp ? tc.get() : nullptr);
            P.apply(ga, comp ? tc.get() : nullptr, fe.get());
            est = device_norm(ctx, fe.get(), n) / r.solution_norm;
        } else if (r.solution_norm == 0.0 && r.normal_residual_norm == 0.0) {
            est = 0.0;
        }
        r.forward_error_estimate = est;
        trace.push_back(make_iteration_record(r, 0, it, 0));
        if (better(r)) {
            best = r;
            have = true;
            copy_block<double, double>(s, n, 1, x_dev, 1, xbest.get(), 1, 1.0);
        }
        report.forward_error_estimate = best.forward_error_estimate;
        copy_block<double, double>(s, m, 1, validator.last_residual_negated(), 1, u.get(), 1, -1.0);
        if (lambda > 0.0) copy_block<double, double>(s, n, 1, x_dev, 1, u.get() + m, 1, -root);
        const bool fwd_ok = !(o.forward_tolerance > 0.0) || (std::isfinite(est) && est <= o.forward_tolerance);
        *reached = r.status == SolverStatus::success && fwd_ok;
        return r;
    };

    bool done = false;
    constexpr int kLongest = 128;
    int total = 0, len = 4, weak = 0, stalled = 0;
    LeastSquaresResult cur = certify(0, &done);
    while (!done && total < max_it) {
        const double before = cur.forward_error_estimate, cert_before = cert_of(cur);
        const int took = cycle(std::min(len, max_it - total));
        total += took;
        report.iterations = total;
        if (took == 0) break;
        cur = certify(total, &done);
        const double after = cur.forward_error_estimate, cert_after = cert_of(cur);
        // A cycle that gained less than a factor 10 gets a successor twice as
        // long (the preconditioned system is not orthonormal yet: a partial
        // basis leaves the complement to LSQR).  Once cycles are at their
        // longest, three in a row without halving the estimate or the
        // certificate end the refinement.
        const bool gained_est = std::isfinite(after) && (!std::isfinite(before) || after < 0.5 * before);
        const bool gained_cert = std::isfinite(cert_after) && (!std::isfinite(cert_before) || cert_after < 0.5 * cert_before);
        if (!(std::isfinite(after) && std::isfinite(before) && after < 0.1 * before)) len = std::min(2 * len, kLongest);
        weak = (gained_est || gained_cert || len < kLongest) ? 0 : weak + 1;
        if (weak >= 3) break;
        // the forward target is met but the certificate has reached its
        // rounding floor: more iterations cannot certify
        const bool fwd_met = o.forward_tolerance > 0.0 && std::isfinite(after) && after <= o.forward_tolerance;
        stalled = fwd_met && !gained_cert ? stalled + 1 : 0;
        if (stalled >= 2) break;
    }
    // the best candidate's x back into x_dev
    copy_block<double, double>(s, n, 1, xbest.get(), 1, x_dev, 1, 1.0);
    ctx.sync();
    if (!done && best.status == SolverStatus::success) {
        // certified, but the forward-error target was not reached
        best.status = SolverStatus::precision_limit;
        best.stop_reason = StopReason::precision_limit;
    }
    return best;
}

}  // namespace amflsx
