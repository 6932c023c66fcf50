// Result, statistics and option types of the AMFLS HIP engine.  Field names,
// enum values and meanings match the reference's public API
// (include/amfls/least_squares_result.hpp:8-229, include/amfls/amfls.hpp:10-20)
// so results translate one-to-one into its Torch result layout.
#pragma once

#include <cstdint>
#include <limits>
#include <vector>

namespace amflsx {

enum class SolverStatus : int {
    success = 0,
    work_limit = 1,
    basis_limit = 2,
    precision_limit = 3,
    numerical_breakdown = 4
};

enum class StopReason : int {
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

enum class RidgeCorrection : int {
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

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kInfD = std::numeric_limits<double>::infinity();

struct RunStatistics {
    long long a_columns = 0, at_columns = 0, a_block_calls = 0, at_block_calls = 0;
    long long search_a_columns = 0, search_at_columns = 0;
    long long validation_a_columns = 0, validation_at_columns = 0;
    long long sketch_a_columns = 0, sketch_at_columns = 0, sketch_a_block_calls = 0,
              sketch_at_block_calls = 0;
    long long iterative_a_columns = 0, iterative_at_columns = 0, iterative_a_block_calls = 0,
              iterative_at_block_calls = 0;
    long long gaussian_random_columns = 0, gaussian_random_values = 0;
    double total_seconds = 0.0, a_seconds = 0.0, at_seconds = 0.0, orthogonalization_seconds = 0.0,
           projected_solve_seconds = 0.0, validation_seconds = 0.0;
    long long search_a_block_calls = 0, search_at_block_calls = 0;
    long long validation_a_block_calls = 0, validation_at_block_calls = 0;
    long long base_validation_a_columns = 0, base_validation_at_columns = 0;
    long long base_validation_a_block_calls = 0, base_validation_at_block_calls = 0;
    long long ridge_correction_a_columns = 0, ridge_correction_at_columns = 0;
    long long ridge_correction_a_block_calls = 0, ridge_correction_at_block_calls = 0;
    double base_validation_seconds = 0.0, ridge_correction_validation_seconds = 0.0;
    long long gaussian_random_block_requests = 0;
};

struct IterationRecord {
    int epoch = 0, depth = 0, auxiliary_width = 0, basis_rank = 0;
    long long a_columns = 0, at_columns = 0, a_block_calls = 0, at_block_calls = 0;
    double objective = 0.0, residual_norm = 0.0, augmented_residual_norm = 0.0,
           normal_residual_norm = 0.0, operator_norm_lower_bound = 0.0;
    double compatible_backward_error_upper_bound = 0.0, least_squares_backward_error_upper_bound = 0.0,
           backward_error_upper_bound = 0.0, relative_energy_error_upper_bound = 0.0,
           relative_normal_residual_upper_bound = 0.0;
    double solution_norm = 0.0, solution_energy_norm = 0.0, augmented_operator_norm_lower_bound = 0.0,
           energy_error_upper_bound = 0.0, objective_gap_upper_bound = 0.0;
    SolverStatus contract_status = SolverStatus::work_limit;
    StopReason contract_stop_reason = StopReason::maximum_depth;
    bool contract_passed = false;
    RidgeCorrection ridge_correction_disposition = RidgeCorrection::not_applicable;
    double ridge_base_energy_error_upper_bound = kNaN, ridge_corrected_energy_error_upper_bound = kNaN,
           ridge_correction_gamma = kNaN, ridge_correction_z_h_z = kNaN,
           ridge_correction_two_abs_z_t_q = kNaN, ridge_correction_q_norm_squared = kNaN;
    // Engine extension: seconds of the search interval that preceded this
    // candidate (operator / local), used to train the width predictor.
    double interval_operator_seconds = 0.0, interval_local_seconds = 0.0;
    int active_width = 0;
};

struct LeastSquaresResult {
    int rows = 0, cols = 0, iterations = 0, depth = 0, auxiliary_width = 0, basis_rank = 0;
    std::vector<double> solution;
    double objective = 0.0, residual_norm = 0.0, augmented_residual_norm = 0.0,
           normal_residual_norm = 0.0, solution_norm = 0.0, solution_energy_norm = 0.0,
           operator_norm_lower_bound = 0.0, augmented_operator_norm_lower_bound = 0.0;
    double compatible_backward_error_upper_bound = 0.0, least_squares_backward_error_upper_bound = 0.0,
           backward_error_upper_bound = 0.0, relative_normal_residual_upper_bound = 0.0;
    double energy_error_upper_bound = 0.0, relative_energy_error_upper_bound = 0.0,
           objective_gap_upper_bound = 0.0;
    SolverStatus status = SolverStatus::work_limit;
    StopReason stop_reason = StopReason::maximum_depth;
    RunStatistics statistics;
    std::vector<IterationRecord> trace;
    RidgeCorrection ridge_correction_disposition = RidgeCorrection::not_applicable;
    double ridge_base_energy_error_upper_bound = kNaN, ridge_corrected_energy_error_upper_bound = kNaN,
           ridge_correction_gamma = kNaN, ridge_correction_z_h_z = kNaN,
           ridge_correction_two_abs_z_t_q = kNaN, ridge_correction_q_norm_squared = kNaN;
    // Engine extensions (mixed precision): LSQR iterations of the FP64
    // refinement, and the forward-error estimate ||P P^T g|| / ||x|| of the
    // returned solution (P the preconditioner from the FP32 search, g the
    // FP64 gradient); NaN where no estimate was made.
    int refinement_steps = 0;
    // mixed precision with a forward-error target: the refinement missed the
    // target and the FP64 engine produced this result
    bool fp64_fallback = false;
    double forward_error_estimate = kNaN;
};

// Arithmetic of the search process.
// fp64: binary64 search; mixed: FP32 search inside an FP64 refinement loop;
// fp32: one FP32 search process with the FP64 certificate (no refinement).
enum class Precision : int { fp64 = 0, mixed = 1, fp32 = 2 };
// 0 = wall clock (hipEvent/steady_clock), 1 = deterministic synthetic clock.
enum class ClockMode : int { wall = 0, synthetic = 1 };

struct SolveOptions {
    double regularization = 0.0;
    double tolerance = 1e-8;
    double failure_probability = 1e-6;
    int maximum_epochs = 0;
    int maximum_basis_size = 0;
    uint64_t seed = 0;
    uint64_t stream = 0;
    int maximum_depth = 0;
    int maximum_auxiliary_width = 0;
    // engine extensions
    Precision precision = Precision::fp64;
    ClockMode clock_mode = ClockMode::wall;
    double synthetic_operator_unit = 1.0e-3;
    double synthetic_local_unit = 1.0e-6;
    int maximum_refinements = 8;
    double inner_tolerance = 0.0;  // 0 = default for mixed mode
    int width_policy = 0;           // 0 = reference heuristic, 1 = learned predictor
    bool stop_on_stagnation = false;  // inner solves of the refinement loop
    // mixed precision: the refinement also runs until the forward-error
    // estimate is at most this (0 = certificate only), for at most this many
    // LSQR iterations (0 = 2000)
    double forward_tolerance = 0.0;
    int maximum_refinement_iterations = 0;
    // rank-cutoff factor of FP32 search blocks (0 = default rule)
    double fp32_rank_factor = 0.0;
};

struct FixedRbgkOptions {
    double regularization = 0.0;
    double tolerance = 1e-8;
    double failure_probability = 1e-6;
    int auxiliary_width = 4;
    int maximum_depth = 32;
    int maximum_basis_size = 0;
    uint64_t seed = 0;
    uint64_t stream = 0;
    bool stop_early = true;
    Precision precision = Precision::fp64;
    ClockMode clock_mode = ClockMode::wall;
};

struct LsqrOptions {
    double regularization = 0.0;
    double tolerance = 1e-8;
    int maximum_iterations = 0;
    ClockMode clock_mode = ClockMode::wall;
};

}  // namespace amflsx
