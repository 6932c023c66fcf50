// Solver drivers of the AMFLS HIP engine.
#pragma once

#include <functional>

#include "operator.hpp"
#include "types.hpp"

namespace amflsx {

// Optional replacement of the reference's widening predicate
// (amfls.cpp:322-358).  Receives the controller features and the heuristic
// decision; returns whether to widen now.  The FP64 certificate remains the
// only termination test.
struct WidthFeatures {
    int active_width = 0;
    int proposed_increment = 0;
    double bound_decrease[4] = {0, 0, 0, 0};  // log10 certificate decreases of the last intervals
    double step_seconds[4] = {0, 0, 0, 0};    // per-level seconds of the last intervals
    double rows = 0, cols = 0, nnz = 0;
    double certificate = 0, tolerance = 0;
    int depth = 0, basis_rank = 0, basis_limit = 0;
    bool heuristic = false;
};
using WidthPolicy = std::function<bool(const WidthFeatures&)>;

// b_dev: device FP64 right-hand side (rows of op).
LeastSquaresResult solve_amfls(const DeviceOperator& op, const double* b_dev, const SolveOptions& opts,
                               Context& ctx, const WidthPolicy* policy = nullptr);
LeastSquaresResult solve_fixed_rbgk(const DeviceOperator& op, const double* b_dev, const FixedRbgkOptions& opts,
                                    Context& ctx);
LeastSquaresResult solve_lsqr(const DeviceOperator& op, const double* b_dev, const LsqrOptions& opts,
                              Context& ctx);

}  // namespace amflsx
