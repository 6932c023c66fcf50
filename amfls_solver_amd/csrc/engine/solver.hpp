// Solver drivers of the AMFLS HIP engine.
#pragma once

#include <functional>
#include <vector>

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

// Regularization sweep (engine extension).  The block Golub-Kahan basis of
// (A, b) does not depend on lambda, so one process serves every value: the
// driver runs for the smallest lambda, and at each of its checkpoints the
// current basis is also evaluated (projected solve + FP64 certificate) for
// every other lambda not yet certified.  A lambda that is still uncertified
// when the driver stops is solved on its own.  Each result is certified for
// its own lambda exactly as a single solve's; statistics and timings of the
// swept results are those of the shared process.  FP64 or FP32 precision.
std::vector<LeastSquaresResult> solve_amfls_sweep(const DeviceOperator& op, const double* b_dev,
                                                  const SolveOptions& opts, const std::vector<double>& lambdas,
                                                  Context& ctx, const WidthPolicy* policy = nullptr);

// Independent problems solved concurrently: `workers` host threads, each
// with its own context and stream on the caller's device, take problems in
// order and run their regularization sweeps.  Operators must not hold
// per-context scratch shared across threads (CSR operators qualify, also
// when one operator serves several problems).
struct BatchProblem {
    const DeviceOperator* op = nullptr;
    const double* b_dev = nullptr;
    std::vector<double> lambdas;
};
std::vector<std::vector<LeastSquaresResult>> solve_amfls_batch(const std::vector<BatchProblem>& problems,
                                                               const SolveOptions& opts, int workers,
                                                               const WidthPolicy* policy = nullptr);
LeastSquaresResult solve_fixed_rbgk(const DeviceOperator& op, const double* b_dev, const FixedRbgkOptions& opts,
                                    Context& ctx);
LeastSquaresResult solve_lsqr(const DeviceOperator& op, const double* b_dev, const LsqrOptions& opts,
                              Context& ctx);

}  // namespace amflsx
