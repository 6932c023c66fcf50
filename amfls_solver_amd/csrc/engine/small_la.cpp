#include "small_la.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace amflsx {

bool cholesky_upper(const HostMat& G, HostMat& R, double* min_ratio) {
    const int n = G.rows;
    R = HostMat(n, n);
    double dmin = std::numeric_limits<double>::infinity(), dmax = 0.0;
    for (int j = 0; j < n; ++j) {
        long double d = G(j, j);
        for (int k = 0; k < j; ++k) d -= static_cast<long double>(R(k, j)) * R(k, j);
        if (!(d > 0.0L) || !std::isfinite(static_cast<double>(d))) {
            if (min_ratio) *min_ratio = 0.0;
            return false;
        }
        const double rjj = std::sqrt(static_cast<double>(d));
        R(j, j) = rjj;
        dmin = std::min(dmin, rjj);
        dmax = std::max(dmax, rjj);
        for (int i = j + 1; i < n; ++i) {
            long double v = G(j, i);
            for (int k = 0; k < j; ++k) v -= static_cast<long double>(R(k, j)) * R(k, i);
            R(j, i) = static_cast<double>(v / rjj);
        }
    }
    if (min_ratio) *min_ratio = n > 0 ? dmin / dmax : 1.0;
    return true;
}

HostMat upper_inverse(const HostMat& R) {
    const int n = R.rows;
    HostMat X(n, n);
    for (int j = n - 1; j >= 0; --j) {
        X(j, j) = 1.0 / R(j, j);
        for (int i = j - 1; i >= 0; --i) {
            long double s = 0.0L;
            for (int k = i + 1; k <= j; ++k) s += static_cast<long double>(R(i, k)) * X(k, j);
            X(i, j) = static_cast<double>(-s / R(i, i));
        }
    }
    return X;
}

HostMat matmul(const HostMat& A, const HostMat& B) {
    HostMat C(A.rows, B.cols);
    for (int i = 0; i < A.rows; ++i)
        for (int k = 0; k < A.cols; ++k) {
            const double a = A(i, k);
            if (a == 0.0) continue;
            for (int j = 0; j < B.cols; ++j) C(i, j) += a * B(k, j);
        }
    return C;
}

namespace {

// Dot products of two contiguous columns with four independent partial
// sums (lets the compiler keep them in vector registers).
inline void column_dots(const double* x, const double* y, int m, double& xx, double& yy, double& xy) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    int i = 0;
    for (; i + 4 <= m; i += 4)
        for (int l = 0; l < 4; ++l) {
            a[l] += x[i + l] * x[i + l];
            b[l] += y[i + l] * y[i + l];
            c[l] += x[i + l] * y[i + l];
        }
    for (; i < m; ++i) {
        a[0] += x[i] * x[i];
        b[0] += y[i] * y[i];
        c[0] += x[i] * y[i];
    }
    xx = (a[0] + a[1]) + (a[2] + a[3]);
    yy = (b[0] + b[1]) + (b[2] + b[3]);
    xy = (c[0] + c[1]) + (c[2] + c[3]);
}

inline void rotate_columns(double* x, double* y, int m, double c, double s) {
    for (int i = 0; i < m; ++i) {
        const double xp = x[i], yq = y[i];
        x[i] = c * xp - s * yq;
        y[i] = s * xp + c * yq;
    }
}

// One-sided Jacobi on the columns of W (m x n, m >= n), accumulating V.
// The sweeps run on column-major copies so every dot product and rotation
// streams two contiguous columns.
void jacobi_columns(HostMat& W, HostMat& V) {
    const int m = W.rows, n = W.cols;
    std::vector<double> wc(static_cast<size_t>(m) * n), vc(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) wc[static_cast<size_t>(j) * m + i] = W(i, j);
    for (int i = 0; i < n; ++i) vc[static_cast<size_t>(i) * n + i] = 1.0;
    const double tol = std::numeric_limits<double>::epsilon() * std::sqrt(static_cast<double>(m));
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p) {
            double* wp = wc.data() + static_cast<size_t>(p) * m;
            double* vp = vc.data() + static_cast<size_t>(p) * n;
            for (int q = p + 1; q < n; ++q) {
                double* wq = wc.data() + static_cast<size_t>(q) * m;
                double alpha, beta, gamma;
                column_dots(wp, wq, m, alpha, beta, gamma);
                if (gamma == 0.0 || std::abs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = std::copysign(1.0, zeta) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                rotate_columns(wp, wq, m, c, s);
                rotate_columns(vp, vc.data() + static_cast<size_t>(q) * n, n, c, s);
            }
        }
        if (!rotated) break;
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) W(i, j) = wc[static_cast<size_t>(j) * m + i];
    V = HostMat(n, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V(i, j) = vc[static_cast<size_t>(j) * n + i];
}

Svd svd_tall(const HostMat& A) {
    const int m = A.rows, n = A.cols;
    HostMat W = A, V;
    jacobi_columns(W, V);
    std::vector<double> norms(n);
    for (int j = 0; j < n; ++j) {
        long double s = 0.0L;
        for (int i = 0; i < m; ++i) s += static_cast<long double>(W(i, j)) * W(i, j);
        norms[j] = std::sqrt(static_cast<double>(s));
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return norms[a] > norms[b]; });
    Svd out;
    out.U = HostMat(m, n);
    out.Vt = HostMat(n, n);
    out.s.resize(n);
    for (int k = 0; k < n; ++k) {
        const int j = order[k];
        out.s[k] = norms[j];
        for (int i = 0; i < m; ++i) out.U(i, k) = norms[j] > 0.0 ? W(i, j) / norms[j] : 0.0;
        for (int i = 0; i < n; ++i) out.Vt(k, i) = V(i, j);
    }
    // Columns for zero singular values: complete to an orthonormal set
    // (their values never enter a result but keep U well formed).
    for (int k = 0; k < n; ++k) {
        if (out.s[k] > 0.0) continue;
        for (int e = 0; e < m; ++e) {
            std::vector<double> v(m, 0.0);
            v[e] = 1.0;
            for (int l = 0; l < n; ++l) {
                if (l == k || (out.s[l] == 0.0 && l > k)) continue;
                double d = 0.0;
                for (int i = 0; i < m; ++i) d += out.U(i, l) * v[i];
                for (int i = 0; i < m; ++i) v[i] -= d * out.U(i, l);
            }
            double nv = 0.0;
            for (double x : v) nv += x * x;
            nv = std::sqrt(nv);
            if (nv > 0.5) {
                for (int i = 0; i < m; ++i) out.U(i, k) = v[i] / nv;
                break;
            }
        }
    }
    return out;
}

}  // namespace

Svd jacobi_svd(const HostMat& A) {
    // Square matrices run on the columns of A^T (the rows of A): for the
    // triangular factors of CholeskyQR and staircase QR, whose rows are
    // graded, this takes about half the sweeps of the columns of A.
    if (A.rows > A.cols) return svd_tall(A);
    HostMat At(A.cols, A.rows);
    for (int i = 0; i < A.rows; ++i)
        for (int j = 0; j < A.cols; ++j) At(j, i) = A(i, j);
    Svd t = svd_tall(At);  // A^T = U s Vt  =>  A = Vt^T s U^T
    Svd out;
    const int p = A.rows;
    out.s = t.s;
    out.U = HostMat(A.rows, p);
    out.Vt = HostMat(p, A.cols);
    for (int i = 0; i < A.rows; ++i)
        for (int k = 0; k < p; ++k) out.U(i, k) = t.Vt(k, i);
    for (int k = 0; k < p; ++k)
        for (int j = 0; j < A.cols; ++j) out.Vt(k, j) = t.U(j, k);
    return out;
}

HostMat Staircase::dense() const {
    HostMat D(rows, cols());
    for (int j = 0; j < cols(); ++j)
        for (size_t t = 0; t < val[j].size(); ++t) D(lo[j] + static_cast<int>(t), j) = val[j][t];
    return D;
}

namespace {

// Householder QR of a staircase matrix applied to the rhs; returns false if
// a diagonal entry is (numerically) absent.  R is stored column-wise in
// `work` (column j valid on rows [top[j], j]).
struct StairQr {
    std::vector<int> top, hi;
    std::vector<std::vector<double>> col;  // dense over [top[j], hi[j])
    std::vector<double> rdiag;
    std::vector<double> c;  // Q^T g
    bool ok = true;
};

StairQr staircase_qr(const Staircase& T, const std::vector<double>& g) {
    const int k = T.cols();
    StairQr q;
    q.top.resize(k);
    q.hi.resize(k);
    q.col.resize(k);
    q.rdiag.assign(k, 0.0);
    q.c = g;
    std::vector<std::vector<double>> v(k);  // reflector j over rows [j, hi[j])
    std::vector<double> tau(k, 0.0);
    for (int j = 0; j < k; ++j) {
        const int lo = T.lo[j];
        const int hi = lo + static_cast<int>(T.val[j].size());
        q.hi[j] = std::max(hi, j + 1);
        // first earlier reflector whose row range reaches this column
        int top = std::min(lo, j);
        {
            int a = 0, b = j;  // smallest i in [0, j) with hi[i] > lo
            while (a < b) {
                const int mid = (a + b) / 2;
                if (q.hi[mid] > lo) b = mid; else a = mid + 1;
            }
            if (a < j) top = std::min(top, a);
        }
        q.top[j] = top;
        std::vector<double>& x = q.col[j];
        x.assign(q.hi[j] - top, 0.0);
        for (size_t t = 0; t < T.val[j].size(); ++t) x[lo + t - top] = T.val[j][t];
        for (int i = top; i < j; ++i) {
            if (tau[i] == 0.0) continue;
            const int end = q.hi[i];
            double d = 0.0;
            for (int r = i; r < end; ++r) d += v[i][r - i] * x[r - top];
            d *= tau[i];
            for (int r = i; r < end; ++r) x[r - top] -= d * v[i][r - i];
        }
        // reflector for rows [j, hi[j])
        const int len = q.hi[j] - j;
        double nrm = 0.0;
        {
            double scale = 0.0, ssq = 1.0;
            for (int r = 0; r < len; ++r) {
                const double a = std::abs(x[j - top + r]);
                if (a == 0.0) continue;
                if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
                else ssq += (a / scale) * (a / scale);
            }
            nrm = scale * std::sqrt(ssq);
        }
        const double alpha = x[j - top];
        v[j].assign(len, 0.0);
        if (nrm == 0.0) {
            tau[j] = 0.0;
            q.rdiag[j] = 0.0;
            continue;
        }
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        tau[j] = (beta - alpha) / beta;
        const double inv = 1.0 / (alpha - beta);
        v[j][0] = 1.0;
        for (int r = 1; r < len; ++r) v[j][r] = x[j - top + r] * inv;
        x[j - top] = beta;
        for (int r = 1; r < len; ++r) x[j - top + r] = 0.0;
        q.rdiag[j] = beta;
        // apply to rhs
        double d = 0.0;
        for (int r = 0; r < len; ++r) d += v[j][r] * q.c[j + r];
        d *= tau[j];
        for (int r = 0; r < len; ++r) q.c[j + r] -= d * v[j][r];
    }
    return q;
}

void back_substitute(const StairQr& q, std::vector<double>& y) {
    const int k = static_cast<int>(q.top.size());
    std::vector<double> c(q.c.begin(), q.c.begin() + k);
    y.assign(k, 0.0);
    for (int j = k - 1; j >= 0; --j) {
        const double yj = c[j] / q.rdiag[j];
        y[j] = yj;
        for (int i = q.top[j]; i < j; ++i) c[i] -= q.col[j][i - q.top[j]] * yj;
    }
}

bool dense_min_norm(const Staircase& T, const std::vector<double>& g, double rcond, std::vector<double>& y) {
    HostMat D = T.dense();
    Svd sv = jacobi_svd(D);
    const int k = T.cols();
    y.assign(k, 0.0);
    if (sv.s.empty() || !(sv.s[0] > 0.0)) return true;
    const double cut = rcond * sv.s[0];
    for (size_t i = 0; i < sv.s.size(); ++i) {
        if (!(sv.s[i] > cut)) break;
        long double d = 0.0L;
        for (int r = 0; r < D.rows; ++r) d += static_cast<long double>(sv.U(r, static_cast<int>(i))) * g[r];
        const double coef = static_cast<double>(d) / sv.s[i];
        for (int j = 0; j < k; ++j) y[j] += sv.Vt(static_cast<int>(i), j) * coef;
    }
    return true;
}

// [T; sqrt(lambda) I] as a staircase: damping row j goes right after the
// last T row of column j, which keeps the staircase shape; g is permuted
// along (damping rows get 0).
void damped_staircase(const Staircase& T, const std::vector<double>& g, double lambda, Staircase& A,
                      std::vector<double>& ga) {
    const int k = T.cols();
    const double root = std::sqrt(lambda);
    std::vector<int> hi(k);
    for (int j = 0; j < k; ++j) hi[j] = T.lo[j] + static_cast<int>(T.val[j].size());
    // new position of T row r: r + #{damping rows placed before it}; damping
    // row j is placed after T row hi[j]-1, i.e. before T row hi[j]
    std::vector<int> pos_t(T.rows + 1, 0);
    {
        int jd = 0;
        for (int r = 0; r <= T.rows; ++r) {
            while (jd < k && hi[jd] <= r) ++jd;  // damping rows with hi <= r come first
            pos_t[r] = r + jd;
        }
    }
    A.rows = T.rows + k;
    A.lo.assign(k, 0);
    A.val.assign(k, {});
    ga.assign(A.rows, 0.0);
    for (int r = 0; r < T.rows && r < static_cast<int>(g.size()); ++r) ga[pos_t[r]] = g[r];
    for (int j = 0; j < k; ++j) {
        // damping row j position: rows placed before it = T rows < hi[j]
        // plus damping rows i < j (all have hi[i] <= hi[j]).
        const int pd = hi[j] + j;
        const int lo_new = T.val[j].empty() ? pd : pos_t[T.lo[j]];
        A.lo[j] = lo_new;
        A.val[j].assign(pd - lo_new + 1, 0.0);
        for (size_t t = 0; t < T.val[j].size(); ++t)
            A.val[j][pos_t[T.lo[j] + static_cast<int>(t)] - lo_new] = T.val[j][t];
        A.val[j][pd - lo_new] = root;
    }
}

}  // namespace

bool projected_solve(const Staircase& T, const std::vector<double>& g, double lambda, double rcond,
                     std::vector<double>& y, bool* used_dense_fallback) {
    if (used_dense_fallback) *used_dense_fallback = false;
    const int k = T.cols();
    y.assign(k, 0.0);
    if (k == 0) return true;
    for (const auto& c : T.val)
        for (double v : c)
            if (!std::isfinite(v)) return false;
    for (double v : g)
        if (!std::isfinite(v)) return false;
    if (lambda > 0.0) {
        Staircase A;
        std::vector<double> ga;
        damped_staircase(T, g, lambda, A, ga);
        StairQr q = staircase_qr(A, ga);
        for (double d : q.rdiag)
            if (d == 0.0) return dense_min_norm(T, g, 0.0, y);
        back_substitute(q, y);
        return true;
    }
    bool staircase = T.rows >= k;
    if (staircase) {
        StairQr q = staircase_qr(T, g);
        double dmax = 0.0, dmin = std::numeric_limits<double>::infinity();
        for (double d : q.rdiag) {
            dmax = std::max(dmax, std::abs(d));
            dmin = std::min(dmin, std::abs(d));
        }
        if (dmax > 0.0 && dmin > 100.0 * rcond * dmax) {
            back_substitute(q, y);
            return true;
        }
    }
    if (used_dense_fallback) *used_dense_fallback = true;
    return dense_min_norm(T, g, rcond, y);
}

int staircase_r(const Staircase& T, double lambda, double floor_ratio, BandUpper& R) {
    const int k = T.cols();
    R.top.assign(k, 0);
    R.col.assign(k, {});
    if (k == 0) return 0;
    std::vector<double> zero(static_cast<size_t>(std::max(T.rows, k) + k), 0.0);
    StairQr q;
    if (lambda > 0.0) {
        Staircase A;
        std::vector<double> ga;
        damped_staircase(T, zero, lambda, A, ga);
        q = staircase_qr(A, ga);
    } else {
        q = staircase_qr(T, zero);
    }
    double dmax = 0.0;
    for (int j = 0; j < k; ++j) dmax = std::max(dmax, std::abs(q.rdiag[j]));
    const double floor = floor_ratio * dmax;
    int raised = 0;
    for (int j = 0; j < k; ++j) {
        R.top[j] = q.top[j];
        R.col[j].assign(q.col[j].begin(), q.col[j].begin() + (j - q.top[j] + 1));
        double& d = R.col[j].back();
        if (!(std::abs(d) >= floor) || d == 0.0) {
            d = d < 0.0 ? -std::max(floor, std::numeric_limits<double>::min())
                        : std::max(floor, std::numeric_limits<double>::min());
            ++raised;
        }
    }
    return raised;
}

void band_solve(const BandUpper& R, std::vector<double>& x) {
    const int k = R.size();
    for (int j = k - 1; j >= 0; --j) {
        const double xj = x[j] / R.diag(j);
        x[j] = xj;
        const std::vector<double>& c = R.col[j];
        const int top = R.top[j];
        for (int i = top; i < j; ++i) x[i] -= c[i - top] * xj;
    }
}

void band_solve_t(const BandUpper& R, std::vector<double>& x) {
    const int k = R.size();
    for (int j = 0; j < k; ++j) {
        const std::vector<double>& c = R.col[j];
        const int top = R.top[j];
        double s = x[j];
        for (int i = top; i < j; ++i) s -= c[i - top] * x[i];
        x[j] = s / R.diag(j);
    }
}

double band_min_singular(const BandUpper& R, int iterations) {
    const int k = R.size();
    if (k == 0) return 0.0;
    std::vector<double> x(k);
    for (int i = 0; i < k; ++i) x[i] = 1.0 + 0.5 * std::sin(1.0 + 0.7 * i);  // fixed, not orthogonal to much
    double lam = 0.0;
    for (int it = 0; it < iterations; ++it) {
        double nx = 0.0;
        for (double v : x) nx += v * v;
        nx = std::sqrt(nx);
        if (!(nx > 0.0) || !std::isfinite(nx)) return 0.0;
        for (double& v : x) v /= nx;
        band_solve_t(R, x);
        band_solve(R, x);  // x <- (R^T R)^{-1} x
        double ny = 0.0;
        for (double v : x) ny += v * v;
        lam = std::sqrt(ny);  // Rayleigh-like growth factor ~ 1 / sigma_min^2
    }
    return lam > 0.0 ? 1.0 / std::sqrt(lam) : 0.0;
}

}  // namespace amflsx
