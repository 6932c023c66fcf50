// Host FP64 small dense linear algebra for the AMFLS engine: Cholesky,
// triangular inverse, one-sided Jacobi SVD and the staircase QR projected
// least-squares solve.
#pragma once

#include <vector>

namespace amflsx {

// Row-major dense matrix helper.
struct HostMat {
    int rows = 0, cols = 0;
    std::vector<double> a;
    HostMat() = default;
    HostMat(int r, int c) : rows(r), cols(c), a(static_cast<size_t>(r) * c, 0.0) {}
    double& operator()(int i, int j) { return a[static_cast<size_t>(i) * cols + j]; }
    double operator()(int i, int j) const { return a[static_cast<size_t>(i) * cols + j]; }
};

// Upper Cholesky factor R with G = R^T R.  Returns false on a nonpositive
// pivot.  `min_ratio` receives min_i R_ii / max_i R_ii.
bool cholesky_upper(const HostMat& G, HostMat& R, double* min_ratio);

// Inverse of an upper triangular matrix.
HostMat upper_inverse(const HostMat& R);

HostMat matmul(const HostMat& A, const HostMat& B);

// Thin SVD A = U diag(s) Vt with singular values in decreasing order
// (one-sided Jacobi; A is rows x cols, any shape).
struct Svd {
    HostMat U;   // rows x p
    std::vector<double> s;  // p = min(rows, cols)
    HostMat Vt;  // p x cols
};
Svd jacobi_svd(const HostMat& A);

// Lower block-bidiagonal ("staircase") projected operator of the block
// Golub-Kahan process: column j holds rows [lo[j], lo[j] + val[j].size()),
// with lo and lo + size nondecreasing in j.
struct Staircase {
    int rows = 0;
    std::vector<int> lo;
    std::vector<std::vector<double>> val;
    int cols() const { return static_cast<int>(lo.size()); }
    HostMat dense() const;
};

// Projected solve of the block Golub-Kahan candidate (reference
// krylov_workspace.cpp:855-901): lambda == 0 gives the minimum-norm least
// squares solution with LAPACK dgelsy's relative rank tolerance `rcond`;
// lambda > 0 gives argmin ||T y - g||^2 + lambda ||y||^2.  Returns false if
// the inputs are not finite.
bool projected_solve(const Staircase& T, const std::vector<double>& g, double lambda, double rcond,
                     std::vector<double>& y, bool* used_dense_fallback);

// Upper triangular k x k factor in column band form: column j holds
// R(top[j] .. j, j) in col[j] (col[j].back() is the diagonal).
struct BandUpper {
    std::vector<int> top;
    std::vector<std::vector<double>> col;
    int size() const { return static_cast<int>(top.size()); }
    double diag(int j) const { return col[j].back(); }
};

// R of the Householder QR of the staircase [T; sqrt(lambda) I] (the
// triangular factor of the projected operator, T^T T + lambda I = R^T R).
// Diagonal entries below `floor_ratio` times the largest are raised to that
// floor (same sign), so R stays invertible on numerically rank-deficient T;
// returns the number of entries raised.
int staircase_r(const Staircase& T, double lambda, double floor_ratio, BandUpper& R);
// x <- R^{-1} x and x <- R^{-T} x (band back / forward substitution).
void band_solve(const BandUpper& R, std::vector<double>& x);
void band_solve_t(const BandUpper& R, std::vector<double>& x);
// Estimate of the smallest singular value of R: `iterations` steps of
// inverse iteration on R^T R from a fixed start vector.
double band_min_singular(const BandUpper& R, int iterations);

}  // namespace amflsx
