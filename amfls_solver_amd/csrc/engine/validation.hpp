// Certified FP64 validation of a candidate on the original problem.
//
// Follows the reference's validate_original_candidate and helpers
// (src/algorithms/mathematics/krylov/candidate_validation.cpp:96-1378): fresh
// FP64 products Ax and A^T(Ax - b), g = fma(lambda, x, A^T r), outward-rounded
// norm intervals, the callback error radii gamma_k ||A|| ||input||, the
// compatible / least-squares backward-error and ridge energy bounds, the
// optional long-double refinement and the one-step ridge correction.  The
// products and reductions run on the device; the bound arithmetic on the host.
#pragma once

#include "operator.hpp"
#include "fp_bounds.hpp"

namespace amflsx {

struct Envelope {
    bool certified = false;
    double operator_norm_upper_bound = kInfD, operator_norm_lower_bound = 0.0;
    double residual_lower = 0.0, residual_upper = kInfD, gradient_upper = kInfD;
    double solution_lower = 0.0, solution_upper = kInfD, right_hand_side_lower = 0.0;
    double applied_solution_lower = 0.0, applied_solution_upper = kInfD, solution_energy_lower = 0.0;
    double gradient_error_radius = kInfD;
    ValidationModel model;
};

class ValidationEngine {
public:
    ValidationEngine(const DeviceOperator& op, const double* b_dev, Context& ctx, RunStatistics& stats);

    // Validate x (device, length cols) for min ||Ax-b||^2 + lambda ||x||^2.
    LeastSquaresResult validate(const double* x_dev, double lambda, double tolerance);

    // Residual r = b - A x of the last validated candidate (device, rows).
    const double* last_residual_negated() const { return r_.get(); }  // holds A x - b
    // Gradient A^T (A x - b) + lambda x of the last validated candidate.
    const double* last_gradient() const { return g_.get(); }
    double rhs_norm() const { return bnorm_; }
    fp::Interval rhs_interval() const { return b_iv_; }

private:
    struct Stats3 { double max = 0, sum = 0, bad = 0; };
    Stats3 stats_of(const double* v, int64_t n);
    // Statistics of several device vectors with one wait.
    void stats_of_many(const double* const* v, const int64_t* n, int count, Stats3* out);
    fp::Interval interval(const Stats3& s, int64_t n) const {
        return fp::norm_interval(s.max, s.sum, s.bad, n);
    }
    double nominal(const Stats3& s) const { return fp::nominal_norm(s.max, s.sum, s.bad); }
    void apply_counted(const double* in, double* out, bool transpose, bool base_phase);
    void count(bool transpose, bool base_phase, double seconds);
    Envelope base_envelope(double lambda, const Stats3& sb, const Stats3& sx, const Stats3& sax,
                           const Stats3& sr, const Stats3& satr, const Stats3& sg);
    bool refine(const double* x_dev, double lambda, Envelope& env, Stats3& sax, Stats3& sr, Stats3& sg,
                double& ax_nominal);

    const DeviceOperator& op_;
    const double* b_dev_;
    Context& ctx_;
    RunStatistics& stats_;
    int64_t m_, n_;
    double bnorm_ = 0.0;
    Stats3 sb_;
    fp::Interval b_iv_;
    DevBuf<double> ax_, r_, atr_, g_, d_, ad_, td_, hd_;
};

void evaluate_accuracy_contract(LeastSquaresResult& res, double bnorm, double lambda, double tol,
                                const double* ridge_bound);
void evaluate_certified_contract(LeastSquaresResult& res, const Envelope& env, double lambda, double tol,
                                 const double* ridge_bound);
bool candidate_validation_failed_numerically(const LeastSquaresResult& r);
IterationRecord make_iteration_record(const LeastSquaresResult& r, int epoch, int depth, int aux_width);

}  // namespace amflsx
