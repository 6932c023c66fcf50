// FP64 refinement of a mixed-precision solve, preconditioned by the FP32
// block Golub-Kahan factorization of the search.
//
// An FP32 search leaves a basis V (n x k, orthonormal to FP32 accuracy) and
// the projected operator T with A V = U T up to FP32 rounding.  With R the
// triangular factor of [T; sqrt(lambda) I] (R^T R = T^T T + lambda I), the
// map P = V R^{-1} makes [A; sqrt(lambda) I] P orthonormal up to about
// kappa(A) * u32, so LSQR (Paige & Saunders, ACM TOMS 8, 1982) on the
// preconditioned correction problem
//     min_z || [A; sqrt(lambda) I] P z - [b - A x0; -sqrt(lambda) x0] ||,
// run entirely in FP64 (products, residual, recurrences), converges in a
// few iterations to the FP64 least-squares solution x = x0 + P z.  This is
// the right-preconditioned LSQR of Blendenpik (Avron, Maymounkov & Toledo,
// SISC 32, 2010) with the preconditioner taken from the low-precision
// Krylov factorization instead of a sketch.  When the basis does not span
// R^n, P adds tau (I - V V^T) on its complement, tau = 1 / sigma_min(R), which
// scales the unexplored directions like the smallest explored one.
//
// Classical refinement that re-solves min ||A d - (b - A x)|| in FP32 keeps
// the full least-squares residual in its right-hand side and stalls at a
// forward error of about kappa^2 u32 ||r|| / (||A|| ||x||) (for the
// 1e6 x 2000 dense problem with kappa = 1e6: ~1e-4); here every quantity
// that carries the residual is FP64.
#pragma once

#include <vector>

#include "operator.hpp"
#include "small_la.hpp"
#include "types.hpp"

namespace amflsx {

// Basis of an FP32 search, copied out before its workspace is released.
struct BasisCapture {
    DevBuf<float> V;      // n x k row-major, leading dimension k
    int k = 0;
    int64_t n = 0;
    Staircase T;
    bool valid = false;
};

// P and P^T on device vectors: a in R^k (host), c in R^n (device, only with
// the complement term), y = P (a, c) in R^n (device).
class BasisPreconditioner {
public:
    BasisPreconditioner(Context& ctx, const BasisCapture& basis, double lambda);
    bool has_complement() const { return complement_; }
    int rank() const { return k_; }
    double tau() const { return tau_; }
    int raised_diagonals() const { return raised_; }
    double sigma_min_r() const { return sigma_min_; }
    // y = P (a, c); c may be null when there is no complement.
    void apply(const std::vector<double>& a, const double* c, double* y);
    // (a, c) = P^T g; waits for the k coefficients.
    void apply_t(const double* g, std::vector<double>& a, double* c);

private:
    Context& ctx_;
    const BasisCapture& basis_;
    int64_t n_;
    int k_;
    bool complement_;
    double tau_ = 1.0, sigma_min_ = 0.0;
    int raised_ = 0;
    BandUpper R_;
    DevBuf<double> proj_;  // k
    std::vector<double> t_;
};

struct RefineReport {
    int iterations = 0;
    int checkpoints = 0;
    double forward_error_estimate = kInfD;
};

// LSQR refinement from x (device, n; updated in place) on the original
// problem min ||A x - b||^2 + lambda ||x||^2.  Every checkpoint certifies x
// with the FP64 validator; the loop ends once the certificate reaches
// o.tolerance and, when o.forward_tolerance > 0, the forward-error estimate
// ||P P^T g|| / ||x|| (g the FP64 gradient; P P^T approximates the inverse
// Hessian) is below it, or after o.maximum_refinement_iterations.  Returns
// the last certified result; products are counted into `stats`.
LeastSquaresResult refine_preconditioned(const DeviceOperator& op, const double* b_dev, const SolveOptions& o,
                                         Context& ctx, RunStatistics& stats, const BasisCapture& basis,
                                         double* x_dev, std::vector<IterationRecord>& trace, RefineReport& report);

}  // namespace amflsx
