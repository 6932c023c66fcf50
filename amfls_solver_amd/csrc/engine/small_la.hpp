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

}  // namespace amflsx
