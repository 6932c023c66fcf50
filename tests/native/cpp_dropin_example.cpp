// A program written against the reference's C++ API (amfls::MatrixOperator,
// amfls::solve_amfls, amfls::LeastSquaresResult), built on the MI355X
// engine's drop-in header.  It prints key=value lines for the test.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "amfls_hip.hpp"

namespace {

// The reference README's matrix-free diagonal operator.
class DiagonalOperator final : public amfls::MatrixOperator {
public:
    explicit DiagonalOperator(std::vector<double> d) : d_(std::move(d)) {}
    int rows() const override { return static_cast<int>(d_.size()); }
    int cols() const override { return static_cast<int>(d_.size()); }
    void apply(const double* x, int s, double* y) const override {
        for (int j = 0; j < s; ++j)
            for (int i = 0; i < rows(); ++i) y[i + j * rows()] = d_[i] * x[i + j * rows()];
    }
    void apply_transpose(const double* y, int s, double* x) const override { apply(y, s, x); }
    amfls::MatrixOperatorValidationErrorModel validation_error_model() const noexcept override {
        double m = 0.0;
        for (double v : d_) m = std::max(m, std::fabs(v));
        return {m, m, 1, 1};
    }

private:
    std::vector<double> d_;
};

// A column-major dense operator with the optional long-double refinement
// products and a counted cost hook.
class DenseOperator final : public amfls::MatrixOperator {
public:
    DenseOperator(int m, int n, std::vector<double> a) : m_(m), n_(n), a_(std::move(a)) {}
    int rows() const override { return m_; }
    int cols() const override { return n_; }
    void apply(const double* x, int s, double* y) const override {
        for (int b = 0; b < s; ++b)
            for (int i = 0; i < m_; ++i) {
                double acc = 0.0;
                for (int j = 0; j < n_; ++j) acc += a_[i + j * m_] * x[j + b * n_];
                y[i + b * m_] = acc;
            }
    }
    void apply_transpose(const double* y, int s, double* x) const override {
        for (int b = 0; b < s; ++b)
            for (int j = 0; j < n_; ++j) {
                double acc = 0.0;
                for (int i = 0; i < m_; ++i) acc += a_[i + j * m_] * y[i + b * m_];
                x[j + b * n_] = acc;
            }
    }
    double relative_block_product_cost(int s) const noexcept override {
        ++cost_queries;
        return s > 0 ? 1.0 : std::numeric_limits<double>::infinity();
    }
    amfls::MatrixOperatorValidationErrorModel validation_error_model() const noexcept override {
        double fro = 0.0, mx = 0.0;
        for (double v : a_) {
            fro += v * v;
            mx = std::max(mx, std::fabs(v));
        }
        return {std::nextafter(std::sqrt(fro) * (1.0 + 1e-15), 1e300), mx, 2LL * n_ + 2, 2LL * m_ + 2};
    }
    bool supports_validation_refinement() const noexcept override { return true; }
    bool apply_validation_refinement(const long double* x, long double* y) const noexcept override {
        ++refinements;
        for (int i = 0; i < m_; ++i) {
            long double acc = 0.0L;
            for (int j = 0; j < n_; ++j) acc += static_cast<long double>(a_[i + j * m_]) * x[j];
            y[i] = acc;
        }
        return true;
    }
    bool apply_transpose_validation_refinement(const long double* y, long double* x) const noexcept override {
        ++refinements;
        for (int j = 0; j < n_; ++j) {
            long double acc = 0.0L;
            for (int i = 0; i < m_; ++i) acc += static_cast<long double>(a_[i + j * m_]) * y[i];
            x[j] = acc;
        }
        return true;
    }
    mutable long long refinements = 0, cost_queries = 0;

private:
    int m_, n_;
    std::vector<double> a_;
};

class ThrowingOperator final : public amfls::MatrixOperator {
public:
    int rows() const override { return 4; }
    int cols() const override { return 3; }
    void apply(const double*, int, double*) const override { throw std::domain_error("operator refused"); }
    void apply_transpose(const double*, int, double*) const override { throw std::domain_error("operator refused"); }
};

}  // namespace

int main() {
    // README example: diag(100, 3, 2, 1) x = (100, 6, 6, 4) -> x = (1, 2, 3, 4)
    DiagonalOperator diag({100.0, 3.0, 2.0, 1.0});
    const double b[] = {100.0, 6.0, 6.0, 4.0};
    amfls::AmflsOptions o;
    o.tolerance = 1e-10;
    o.seed = 7;
    amfls::LeastSquaresResult r = amfls::solve_amfls(diag, b, o);
    double err = 0.0;
    for (int i = 0; i < 4; ++i) err = std::max(err, std::fabs(r.solution[i] - (i + 1)));
    std::printf("diag_status=%d\ndiag_max_error=%.3e\ndiag_trace_rows=%zu\ndiag_backward=%.3e\n",
                static_cast<int>(r.status), err, r.trace.size(), r.backward_error_upper_bound);

    // 6 x 3 integer least-squares problem; an unreachable tolerance makes the
    // validator take the long-double branch through the operator's hooks
    const int m = 6, n = 3;
    const double rowmajor[m][n] = {{1, 2, 0}, {0, 1, 1}, {3, 0, 1}, {1, 1, 1}, {2, -1, 0}, {0, 0, 4}};
    std::vector<double> a(m * n);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) a[i + j * m] = rowmajor[i][j];
    DenseOperator dense(m, n, a);
    const double bd[] = {1, -2, 3, 0, 5, 1};
    amfls::AmflsOptions o2;
    o2.tolerance = 1e-30;
    amfls::LeastSquaresResult r2 = amfls::solve_amfls(dense, bd, o2);
    std::printf("dense_status=%d\ndense_stop=%d\ndense_basis_rank=%d\ndense_refinement_calls=%lld\n",
                static_cast<int>(r2.status), static_cast<int>(r2.stop_reason), r2.basis_rank, dense.refinements);
    std::printf("dense_cost_queries=%lld\ndense_ls_backward=%.3e\ndense_base_validation_a=%lld\n", dense.cost_queries,
                r2.least_squares_backward_error_upper_bound, r2.statistics.base_validation_a_columns);
    for (int j = 0; j < n; ++j) std::printf("dense_x%d=%.17g\n", j, r2.solution[j]);

    // the caller's exception propagates unchanged; invalid options are invalid_argument
    try {
        ThrowingOperator t;
        const double bt[] = {1, 2, 3, 4};
        amfls::solve_amfls(t, bt, amfls::AmflsOptions{});
        std::printf("throwing=none\n");
    } catch (const std::domain_error& e) {
        std::printf("throwing=domain_error:%s\n", e.what());
    }
    try {
        amfls::AmflsOptions bad;
        bad.tolerance = 0.0;
        amfls::solve_amfls(diag, b, bad);
        std::printf("bad_options=none\n");
    } catch (const std::invalid_argument&) {
        std::printf("bad_options=invalid_argument\n");
    }
    return 0;
}
