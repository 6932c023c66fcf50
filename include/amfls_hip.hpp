// C++ drop-in for the reference's amfls::MatrixOperator / solve_amfls API on
// the MI355X engine (header only, over the C ABI of amfls_hip.h; link with
// libamfls_hip.so).
//
// A program written against the reference's headers
//   include/amfls/matrix_operator.hpp:29-89   (MatrixOperator,
//                                             MatrixOperatorValidationErrorModel)
//   include/amfls/amfls.hpp:10-25             (AmflsOptions, solve_amfls)
//   include/amfls/least_squares_result.hpp    (SolverStatus, StopReason,
//                                             RidgeCorrectionDisposition,
//                                             RunStatistics, IterationRecord,
//                                             LeastSquaresResult)
// compiles against this header unchanged: the same names, members, defaults
// and virtual functions.  solve_amfls runs the search on the GPU; the
// caller's operator is called back on the host with column-major FP64
// blocks exactly as in the reference (apply / apply_transpose, the
// scheduler's relative_block_product_cost, the validation error model and
// the optional long-double refinement products), sequentially on the
// calling thread.
//
// Errors are the reference's exceptions: std::invalid_argument for bad
// inputs or options, std::logic_error, std::length_error, std::runtime_error;
// an exception thrown by the caller's operator propagates out of
// solve_amfls unchanged.  Nonfinite operator output is not an error: it ends
// the solve with SolverStatus::numerical_breakdown.
//
// Each thread that calls solve_amfls holds one engine context (its GPU
// stream and workspaces), created on first use and kept until the thread
// exits.
#pragma once

#include <cstdint>
#include <exception>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "amfls_hip.h"

namespace amfls {

struct MatrixOperatorValidationErrorModel {
    double operator_norm_upper_bound = std::numeric_limits<double>::infinity();
    double operator_norm_lower_bound = 0.0;
    long long apply_rounding_steps = 0;
    long long apply_transpose_rounding_steps = 0;
};

class MatrixOperator {
public:
    virtual ~MatrixOperator() = default;
    virtual int rows() const = 0;
    virtual int cols() const = 0;
    // X: cols() x block_cols, Y: rows() x block_cols, column-major.
    virtual void apply(const double* x, int block_cols, double* y) const = 0;
    virtual void apply_transpose(const double* y, int block_cols, double* x) const = 0;
    virtual double relative_block_product_cost(int block_cols) const noexcept {
        return block_cols > 0 ? static_cast<double>(block_cols) : std::numeric_limits<double>::infinity();
    }
    virtual MatrixOperatorValidationErrorModel validation_error_model() const noexcept { return {}; }
    virtual bool supports_validation_refinement() const noexcept { return false; }
    virtual bool apply_validation_refinement(const long double* x, long double* y) const noexcept {
        (void)x;
        (void)y;
        return false;
    }
    virtual bool apply_transpose_validation_refinement(const long double* y, long double* x) const noexcept {
        (void)y;
        (void)x;
        return false;
    }
};

enum class SolverStatus { success = 0, work_limit = 1, basis_limit = 2, precision_limit = 3, numerical_breakdown = 4 };

enum class StopReason {
    certified_optimality = 0,
    exact_stationarity = 1,
    exhausted_search_space = 2,
    maximum_depth = 3,
    maximum_epochs = 4,
    maximum_basis = 5,
    precision_limit = 6,
    numerical_breakdown = 7,
    compatible_backward_error = 8,
    least_squares_backward_error = 9,
    relative_energy_error = 10
};

enum class RidgeCorrectionDisposition : int {
    not_applicable = 0,
    skipped_base_pass = 1,
    skipped_zero_energy = 2,
    skipped_nonfinite_candidate = 3,
    attempted_improved = 4,
    attempted_not_improved = 5,
    attempted_invalid_curvature = 6,
    attempted_partial_failure = 7,
    attempted_nonfinite = 8
};

struct RunStatistics {
    long long a_columns = 0, at_columns = 0, a_block_calls = 0, at_block_calls = 0;
    long long search_a_columns = 0, search_at_columns = 0, validation_a_columns = 0, validation_at_columns = 0;
    long long sketch_a_columns = 0, sketch_at_columns = 0, sketch_a_block_calls = 0, sketch_at_block_calls = 0;
    long long iterative_a_columns = 0, iterative_at_columns = 0, iterative_a_block_calls = 0,
              iterative_at_block_calls = 0;
    long long gaussian_random_columns = 0, gaussian_random_values = 0;
    double total_seconds = 0.0, a_seconds = 0.0, at_seconds = 0.0, orthogonalization_seconds = 0.0,
           projected_solve_seconds = 0.0, validation_seconds = 0.0;
    long long search_a_block_calls = 0, search_at_block_calls = 0, validation_a_block_calls = 0,
              validation_at_block_calls = 0;
    long long base_validation_a_columns = 0, base_validation_at_columns = 0, base_validation_a_block_calls = 0,
              base_validation_at_block_calls = 0;
    long long ridge_correction_a_columns = 0, ridge_correction_at_columns = 0, ridge_correction_a_block_calls = 0,
              ridge_correction_at_block_calls = 0;
    double base_validation_seconds = 0.0, ridge_correction_validation_seconds = 0.0;
    long long gaussian_random_block_requests = 0;
};

struct IterationRecord {
    int epoch = 0, depth = 0, auxiliary_width = 0, basis_rank = 0;
    long long a_columns = 0, at_columns = 0, a_block_calls = 0, at_block_calls = 0;
    double objective = 0.0, residual_norm = 0.0, augmented_residual_norm = 0.0, normal_residual_norm = 0.0;
    double operator_norm_lower_bound = 0.0;
    double compatible_backward_error_upper_bound = 0.0, least_squares_backward_error_upper_bound = 0.0;
    double backward_error_upper_bound = 0.0, relative_energy_error_upper_bound = 0.0;
    double relative_normal_residual_upper_bound = 0.0;
    double solution_norm = 0.0, solution_energy_norm = 0.0, augmented_operator_norm_lower_bound = 0.0;
    double energy_error_upper_bound = 0.0, objective_gap_upper_bound = 0.0;
    SolverStatus contract_status = SolverStatus::work_limit;
    StopReason contract_stop_reason = StopReason::maximum_depth;
    bool contract_passed = false;
    RidgeCorrectionDisposition ridge_correction_disposition = RidgeCorrectionDisposition::not_applicable;
    double ridge_base_energy_error_upper_bound = std::numeric_limits<double>::quiet_NaN();
    double ridge_corrected_energy_error_upper_bound = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_gamma = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_z_h_z = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_two_abs_z_t_q = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_q_norm_squared = std::numeric_limits<double>::quiet_NaN();
};

struct LeastSquaresResult {
    int rows = 0, cols = 0, iterations = 0, depth = 0, auxiliary_width = 0, basis_rank = 0;
    std::vector<double> solution;
    double objective = 0.0, residual_norm = 0.0, augmented_residual_norm = 0.0, normal_residual_norm = 0.0;
    double solution_norm = 0.0, solution_energy_norm = 0.0, operator_norm_lower_bound = 0.0;
    double augmented_operator_norm_lower_bound = 0.0;
    double compatible_backward_error_upper_bound = 0.0, least_squares_backward_error_upper_bound = 0.0;
    double backward_error_upper_bound = 0.0;
    double relative_normal_residual_upper_bound = 0.0;
    double energy_error_upper_bound = 0.0, relative_energy_error_upper_bound = 0.0, objective_gap_upper_bound = 0.0;
    SolverStatus status = SolverStatus::work_limit;
    StopReason stop_reason = StopReason::maximum_depth;
    RunStatistics statistics;
    std::vector<IterationRecord> trace;
    RidgeCorrectionDisposition ridge_correction_disposition = RidgeCorrectionDisposition::not_applicable;
    double ridge_base_energy_error_upper_bound = std::numeric_limits<double>::quiet_NaN();
    double ridge_corrected_energy_error_upper_bound = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_gamma = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_z_h_z = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_two_abs_z_t_q = std::numeric_limits<double>::quiet_NaN();
    double ridge_correction_q_norm_squared = std::numeric_limits<double>::quiet_NaN();
};

struct AmflsOptions {
    double regularization = 0.0;
    double tolerance = 1e-8;
    double failure_probability = 1e-6;
    int maximum_epochs = 0;
    int maximum_basis_size = 0;
    std::uint64_t seed = 0;
    std::uint64_t stream = 0;
    int maximum_depth = 0;
    int maximum_auxiliary_width = 0;
};

namespace hip_detail {

// Rethrows a failed C call as the exception the reference throws.
inline void check(int code) {
    if (code == AMFLS_OK) return;
    const char* m = amfls_last_error();
    const std::string msg = m ? m : "amfls engine error";
    switch (code) {
        case AMFLS_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case AMFLS_ERR_LOGIC: throw std::logic_error(msg);
        case AMFLS_ERR_LENGTH: throw std::length_error(msg);
        default: throw std::runtime_error(msg);
    }
}

struct ContextHandle {
    amfls_context* ctx = nullptr;
    ContextHandle() { check(amfls_context_create(&ctx)); }
    ~ContextHandle() {
        if (ctx) amfls_context_destroy(ctx);
    }
    ContextHandle(const ContextHandle&) = delete;
    ContextHandle& operator=(const ContextHandle&) = delete;
};

inline amfls_context* thread_context() {
    thread_local ContextHandle handle;
    return handle.ctx;
}

// The caller's operator behind the C callbacks; the first exception a
// callback raises is kept and rethrown once the engine has returned.
struct Bridge {
    const MatrixOperator* op;
    std::exception_ptr error;

    template <typename F>
    int guarded(F&& f) noexcept {
        if (error) return 1;
        try {
            f();
            return 0;
        } catch (...) {
            error = std::current_exception();
            return 1;
        }
    }
    static int apply(void* user, const double* in, int s, double* out) {
        auto* b = static_cast<Bridge*>(user);
        return b->guarded([&] { b->op->apply(in, s, out); });
    }
    static int apply_transpose(void* user, const double* in, int s, double* out) {
        auto* b = static_cast<Bridge*>(user);
        return b->guarded([&] { b->op->apply_transpose(in, s, out); });
    }
    static double cost(void* user, int s) {
        return static_cast<Bridge*>(user)->op->relative_block_product_cost(s);
    }
    static int refine(void* user, const long double* in, long double* out) {
        return static_cast<Bridge*>(user)->op->apply_validation_refinement(in, out) ? 0 : 1;
    }
    static int refine_transpose(void* user, const long double* in, long double* out) {
        return static_cast<Bridge*>(user)->op->apply_transpose_validation_refinement(in, out) ? 0 : 1;
    }
};

struct OperatorHandle {
    amfls_operator* op = nullptr;
    ~OperatorHandle() {
        if (op) amfls_operator_destroy(op);
    }
};

struct ResultHandle {
    amfls_result* r = nullptr;
    ~ResultHandle() {
        if (r) amfls_result_destroy(r);
    }
};

inline LeastSquaresResult decode(const amfls_result* res, int rows, int cols) {
    int64_t n = 0, trace_rows = 0;
    check(amfls_result_sizes(res, &n, &trace_rows));
    LeastSquaresResult r;
    r.rows = rows;
    r.cols = cols;
    r.solution.assign(static_cast<size_t>(n), 0.0);
    double m[AMFLS_METRICS_WIDTH], t[AMFLS_TIMINGS_WIDTH];
    int64_t c[AMFLS_COUNTERS_WIDTH], status = 0, stop = 0;
    std::vector<double> tf(static_cast<size_t>(trace_rows) * AMFLS_TRACE_FLOAT_WIDTH);
    std::vector<int64_t> ti(static_cast<size_t>(trace_rows) * AMFLS_TRACE_INT_WIDTH);
    check(amfls_result_copy(res, r.solution.data(), m, c, t, tf.data(), ti.data(), &status, &stop));
    r.objective = m[0];
    r.residual_norm = m[1];
    r.augmented_residual_norm = m[2];
    r.normal_residual_norm = m[3];
    r.solution_norm = m[4];
    r.solution_energy_norm = m[5];
    r.operator_norm_lower_bound = m[6];
    r.augmented_operator_norm_lower_bound = m[7];
    r.compatible_backward_error_upper_bound = m[8];
    r.least_squares_backward_error_upper_bound = m[9];
    r.backward_error_upper_bound = m[10];
    r.relative_normal_residual_upper_bound = m[11];
    r.energy_error_upper_bound = m[12];
    r.relative_energy_error_upper_bound = m[13];
    r.objective_gap_upper_bound = m[14];
    r.ridge_base_energy_error_upper_bound = m[15];
    r.ridge_corrected_energy_error_upper_bound = m[16];
    r.ridge_correction_gamma = m[17];
    r.ridge_correction_z_h_z = m[18];
    r.ridge_correction_two_abs_z_t_q = m[19];
    r.ridge_correction_q_norm_squared = m[20];
    RunStatistics& s = r.statistics;
    long long* const counters[] = {
        &s.a_columns, &s.at_columns, &s.a_block_calls, &s.at_block_calls, &s.search_a_columns,
        &s.search_at_columns, &s.validation_a_columns, &s.validation_at_columns, nullptr, nullptr, nullptr, nullptr,
        &s.sketch_a_columns, &s.sketch_at_columns, &s.sketch_a_block_calls, &s.sketch_at_block_calls,
        &s.iterative_a_columns, &s.iterative_at_columns, &s.iterative_a_block_calls, &s.iterative_at_block_calls,
        &s.gaussian_random_columns, &s.gaussian_random_values, &s.search_a_block_calls, &s.search_at_block_calls,
        &s.validation_a_block_calls, &s.validation_at_block_calls, &s.base_validation_a_columns,
        &s.base_validation_at_columns, &s.base_validation_a_block_calls, &s.base_validation_at_block_calls,
        &s.ridge_correction_a_columns, &s.ridge_correction_at_columns, &s.ridge_correction_a_block_calls,
        &s.ridge_correction_at_block_calls, &s.gaussian_random_block_requests, nullptr};
    static_assert(sizeof(counters) / sizeof(counters[0]) == AMFLS_COUNTERS_WIDTH, "counter layout");
    for (int i = 0; i < AMFLS_COUNTERS_WIDTH; ++i)
        if (counters[i]) *counters[i] = c[i];
    r.iterations = static_cast<int>(c[8]);
    r.depth = static_cast<int>(c[9]);
    r.auxiliary_width = static_cast<int>(c[10]);
    r.basis_rank = static_cast<int>(c[11]);
    r.ridge_correction_disposition = static_cast<RidgeCorrectionDisposition>(c[35]);
    s.total_seconds = t[0];
    s.a_seconds = t[1];
    s.at_seconds = t[2];
    s.orthogonalization_seconds = t[3];
    s.projected_solve_seconds = t[4];
    s.validation_seconds = t[5];
    s.base_validation_seconds = t[7];
    s.ridge_correction_validation_seconds = t[8];
    r.status = static_cast<SolverStatus>(status);
    r.stop_reason = static_cast<StopReason>(stop);
    r.trace.resize(static_cast<size_t>(trace_rows));
    for (int64_t i = 0; i < trace_rows; ++i) {
        const double* f = tf.data() + i * AMFLS_TRACE_FLOAT_WIDTH;
        const int64_t* q = ti.data() + i * AMFLS_TRACE_INT_WIDTH;
        IterationRecord& e = r.trace[static_cast<size_t>(i)];
        e.epoch = static_cast<int>(f[0]);
        e.depth = static_cast<int>(f[1]);
        e.auxiliary_width = static_cast<int>(f[2]);
        e.basis_rank = static_cast<int>(f[3]);
        e.a_columns = static_cast<long long>(f[4]);
        e.at_columns = static_cast<long long>(f[5]);
        e.a_block_calls = static_cast<long long>(f[6]);
        e.at_block_calls = static_cast<long long>(f[7]);
        e.objective = f[8];
        e.residual_norm = f[9];
        e.augmented_residual_norm = f[10];
        e.normal_residual_norm = f[11];
        e.operator_norm_lower_bound = f[12];
        e.compatible_backward_error_upper_bound = f[13];
        e.least_squares_backward_error_upper_bound = f[14];
        e.backward_error_upper_bound = f[15];
        e.relative_normal_residual_upper_bound = f[16];
        e.relative_energy_error_upper_bound = f[17];
        e.solution_norm = f[18];
        e.solution_energy_norm = f[19];
        e.augmented_operator_norm_lower_bound = f[20];
        e.energy_error_upper_bound = f[21];
        e.objective_gap_upper_bound = f[22];
        e.ridge_base_energy_error_upper_bound = f[23];
        e.ridge_corrected_energy_error_upper_bound = f[24];
        e.ridge_correction_gamma = f[25];
        e.ridge_correction_z_h_z = f[26];
        e.ridge_correction_two_abs_z_t_q = f[27];
        e.ridge_correction_q_norm_squared = f[28];
        e.contract_status = static_cast<SolverStatus>(q[0]);
        e.contract_stop_reason = static_cast<StopReason>(q[1]);
        e.contract_passed = q[2] != 0;
        e.ridge_correction_disposition = static_cast<RidgeCorrectionDisposition>(q[3]);
    }
    return r;
}

}  // namespace hip_detail

// reference amfls.hpp:22-25 / amfls.cpp:57-661 on the GPU engine.
inline LeastSquaresResult solve_amfls(const MatrixOperator& matrix, const double* right_hand_side,
                                      const AmflsOptions& options) {
    using namespace hip_detail;
    if (right_hand_side == nullptr) throw std::invalid_argument("right-hand side cannot be null");
    const int rows = matrix.rows(), cols = matrix.cols();
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("operator dimensions must be positive");
    amfls_context* ctx = thread_context();
    Bridge bridge{&matrix, nullptr};
    const MatrixOperatorValidationErrorModel model = matrix.validation_error_model();
    amfls_host_callbacks cb{};
    cb.apply = &Bridge::apply;
    cb.apply_transpose = &Bridge::apply_transpose;
    cb.relative_block_product_cost = &Bridge::cost;
    if (matrix.supports_validation_refinement()) {
        cb.apply_refinement = &Bridge::refine;
        cb.apply_transpose_refinement = &Bridge::refine_transpose;
    }
    cb.user = &bridge;
    OperatorHandle op;
    check(amfls_host_operator_create_ex(ctx, rows, cols, &cb, model.operator_norm_upper_bound,
                                        model.operator_norm_lower_bound, model.apply_rounding_steps,
                                        model.apply_transpose_rounding_steps, &op.op));
    amfls_options o;
    amfls_options_default(&o);
    o.regularization = options.regularization;
    o.tolerance = options.tolerance;
    o.failure_probability = options.failure_probability;
    o.maximum_epochs = options.maximum_epochs;
    o.maximum_basis_size = options.maximum_basis_size;
    o.seed = options.seed;
    o.stream = options.stream;
    o.maximum_depth = options.maximum_depth;
    o.maximum_auxiliary_width = options.maximum_auxiliary_width;
    ResultHandle res;
    const int code = amfls_solve(ctx, op.op, right_hand_side, 0, &o, &res.r);
    if (bridge.error) std::rethrow_exception(bridge.error);
    check(code);
    return decode(res.r, rows, cols);
}

}  // namespace amfls
