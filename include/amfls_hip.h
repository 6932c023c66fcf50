/*
 * C ABI of the MI355X AMFLS engine (libamfls_hip.so).
 *
 * Plain pointers and sizes only; every function returns an AMFLS_* status
 * code and never lets an exception cross the boundary
 * (amfls_last_error() holds the message of the last failure on the calling
 * thread).  Matrix and vector pointers are HIP device pointers unless a
 * parameter says otherwise.
 *
 * Reference interfaces replaced (paths in doloMing/amfls_solver):
 *   amfls_csr_create / amfls_dense_create
 *       -> math::SparseCsrOperator ctor (src/algorithms/mathematics/operators/
 *          sparse_csr_operator.cpp:14-130), math::DenseOperator ctor
 *          (dense_operator.cpp:20-95), math::RowMajorDenseOperator ctor
 *          (row_major_dense_operator.cpp:23-101); the Torch binding builds
 *          these from tensors in TorchMatrixOperator (bindings/torch/
 *          torch_binding.cpp:69-114).
 *   amfls_operator_apply / amfls_operator_apply_transpose
 *       -> MatrixOperator::apply / apply_transpose (include/amfls/
 *          matrix_operator.hpp:36-44), column-major blocks.
 *   amfls_operator_relative_block_product_cost
 *       -> MatrixOperator::relative_block_product_cost (matrix_operator.hpp:50-55)
 *   amfls_operator_validation_model
 *       -> MatrixOperator::validation_error_model (matrix_operator.hpp:57-60)
 *   amfls_solve            -> amfls::solve_amfls (include/amfls/amfls.hpp:22-25)
 *   amfls_solve_fixed_rbgk -> amfls::solve_fixed_rbgk (include/amfls/fixed_rbgk.hpp:22-25)
 *   amfls_solve_lsqr       -> amfls::solve_lsqr (include/amfls/lsqr.hpp:14-17)
 *   amfls_result_copy      -> torch_binding::result_tensors (bindings/torch/
 *          result_tensors.hpp:12-156): metrics 21, counters 36, timings 9,
 *          trace_float 29 and trace_int 4 columns, status, stop reason.
 */
#ifndef AMFLS_HIP_H
#define AMFLS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AMFLS_OK = 0,
    AMFLS_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    AMFLS_ERR_LOGIC = 2,            /* std::logic_error */
    AMFLS_ERR_LENGTH = 3,           /* std::length_error / overflow */
    AMFLS_ERR_RUNTIME = 4,          /* std::runtime_error */
    AMFLS_ERR_HIP = 5,              /* HIP runtime failure */
    AMFLS_ERR_UNKNOWN = 6
};

enum { AMFLS_F64 = 0, AMFLS_F32 = 1, AMFLS_BF16 = 2 };
enum { AMFLS_PRECISION_FP64 = 0, AMFLS_PRECISION_MIXED = 1, AMFLS_PRECISION_FP32 = 2 };
enum { AMFLS_CLOCK_WALL = 0, AMFLS_CLOCK_SYNTHETIC = 1 };
/* Widening predicate: the reference heuristic, the learned model, the
 * heuristic with the model inputs of each decision recorded, or a fixed
 * answer when a stage could widen: always (while the widened block has at
 * most 64 columns) or never. */
enum {
    AMFLS_WIDTH_HEURISTIC = 0,
    AMFLS_WIDTH_LEARNED = 1,
    AMFLS_WIDTH_RECORD = 2,
    AMFLS_WIDTH_WIDEN = 3,
    AMFLS_WIDTH_HOLD = 4,
    AMFLS_WIDTH_RECORD_WIDEN = 5 /* record while always widening */
};

#define AMFLS_METRICS_WIDTH 21
#define AMFLS_COUNTERS_WIDTH 36
#define AMFLS_TIMINGS_WIDTH 9
#define AMFLS_TRACE_FLOAT_WIDTH 29
#define AMFLS_TRACE_INT_WIDTH 4
#define AMFLS_TRACE_EXT_WIDTH 3

#define AMFLS_WIDTH_INPUTS 12
#define AMFLS_WIDTH_PARAMS 5057

typedef struct amfls_context amfls_context;
typedef struct amfls_width_model amfls_width_model;
typedef struct amfls_operator amfls_operator;
typedef struct amfls_result amfls_result;

/* Mirrors amfls::AmflsOptions (include/amfls/amfls.hpp:10-20) plus engine
 * options: precision, clock mode, refinement limits and the width policy. */
typedef struct amfls_options {
    double regularization;
    double tolerance;
    double failure_probability;
    int32_t maximum_epochs;
    int32_t maximum_basis_size;
    uint64_t seed;
    uint64_t stream;
    int32_t maximum_depth;
    int32_t maximum_auxiliary_width;
    int32_t precision;
    int32_t clock_mode;
    double synthetic_operator_unit;
    double synthetic_local_unit;
    int32_t maximum_refinements;
    double inner_tolerance;
    int32_t width_policy;
    /* Mixed precision (engine extension): after the FP32 search, FP64 LSQR
     * preconditioned by the search's factorization refines the solution
     * until the certificate reaches `tolerance` and, when
     * forward_tolerance > 0, the forward-error estimate ||P P^T g|| / ||x||
     * is at most forward_tolerance; at most maximum_refinement_iterations
     * LSQR iterations (0 = 2000).  maximum_refinements = 0 skips the
     * refinement; inner_tolerance > 0 is the FP32 search's tolerance. */
    double forward_tolerance;
    int32_t maximum_refinement_iterations;
    /* Scripted width decisions (engine extension, single solves): the i-th
     * widening decision of the solve takes width_script[i] (0 hold, 1 widen,
     * any other value: the width policy) for i < width_script_length; later
     * decisions follow the width policy.  Replaying a recorded run's choices
     * and forcing one decision either way gives the counterfactual labels of
     * the learned width controller.  NULL / 0: no script. */
    const int8_t* width_script;
    int32_t width_script_length;
} amfls_options;

const char* amfls_last_error(void);
int amfls_version(void);
void amfls_options_default(amfls_options* options);

int amfls_context_create(amfls_context** out);
int amfls_context_destroy(amfls_context* ctx);
int amfls_context_synchronize(amfls_context* ctx);

int amfls_csr_create(amfls_context* ctx, int64_t rows, int64_t cols, int64_t nnz, const int64_t* row_ptr,
                     const int32_t* col_idx, const void* values, int value_type, amfls_operator** out);
int amfls_dense_create(amfls_context* ctx, int64_t rows, int64_t cols, const void* values, int value_type,
                       int row_major, int64_t ld, amfls_operator** out);
/* A caller's operator on the host: the adapter for custom
 * amfls::MatrixOperator subclasses (reference matrix_operator.hpp:29-89).
 * apply(user, x, s, y) computes y = A x and apply_transpose(user, y, s, x)
 * x = A^T y on host column-major FP64 blocks (x: cols x s, y: rows x s,
 * leading dimension = rows of the block, matrix_operator.hpp:36-44),
 * returning 0 on success.  Each product copies through page-locked host
 * memory and waits for the callback.  The validation error model is the
 * caller's (matrix_operator.hpp:57-60: norm_upper = +inf disables
 * certification); relative_block_product_cost is s, the reference default. */
typedef int (*amfls_host_apply_fn)(void* user, const double* in, int block_cols, double* out);
int amfls_host_operator_create(amfls_context* ctx, int64_t rows, int64_t cols, amfls_host_apply_fn apply,
                               amfls_host_apply_fn apply_transpose, void* user, double norm_upper,
                               double norm_lower, int64_t steps_apply, int64_t steps_transpose,
                               amfls_operator** out);
/* The full adapter for amfls::MatrixOperator subclasses
 * (matrix_operator.hpp:29-89): the two products as above, plus
 * relative_block_product_cost(user, s) for the width scheduler (NULL: s,
 * the reference default, matrix_operator.hpp:50-55) and the long-double
 * validation refinement products (matrix_operator.hpp:70-88; both NULL: the
 * operator does not support refinement), which return 0 on success. */
typedef double (*amfls_host_cost_fn)(void* user, int block_cols);
typedef int (*amfls_host_refine_fn)(void* user, const long double* in, long double* out);
typedef struct amfls_host_callbacks {
    amfls_host_apply_fn apply;
    amfls_host_apply_fn apply_transpose;
    amfls_host_cost_fn relative_block_product_cost;
    amfls_host_refine_fn apply_refinement;
    amfls_host_refine_fn apply_transpose_refinement;
    void* user;
} amfls_host_callbacks;
int amfls_host_operator_create_ex(amfls_context* ctx, int64_t rows, int64_t cols, const amfls_host_callbacks* cb,
                                  double norm_upper, double norm_lower, int64_t steps_apply, int64_t steps_transpose,
                                  amfls_operator** out);
int amfls_operator_destroy(amfls_operator* op);
int amfls_operator_shape(const amfls_operator* op, int64_t* rows, int64_t* cols);
int amfls_operator_validation_model(const amfls_operator* op, double* norm_upper, double* norm_lower,
                                    int64_t* apply_steps, int64_t* transpose_steps);
double amfls_operator_relative_block_product_cost(const amfls_operator* op, int block_cols);
/* Y = A X (transpose = 0) or X = A^T Y (transpose = 1) with column-major
 * blocks exactly as MatrixOperator (X cols x s, Y rows x s); dtype
 * AMFLS_F64 or AMFLS_F32 for the blocks. */
int amfls_operator_apply(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                         int block_cols, void* out, int dtype);
/* Same product on the engine's row-major blocks with leading dimensions. */
int amfls_operator_apply_rows(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                              int block_cols, int64_t ld_in, void* out, int64_t ld_out, int dtype);
/* Allocation-free variant of amfls_operator_apply_rows; it has no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under stream capture, serving loops, harnesses that
 * forbid hipMalloc).  The same product on the same blocks, but its only
 * device memory beyond in/out is a file-scope array of 16 MiB per GPU:
 * packed block planes and per-chunk partial sums, fewer chunks when the
 * default would not fit.  No timing events are recorded.  Device CSR and
 * dense operators only; a product that cannot fit the array fails with
 * AMFLS_ERR_LENGTH.  Calls that overlap on one GPU share the array and
 * race; results are independent of earlier calls (everything read is
 * written by the same call). */
int amfls_operator_apply_rows_noalloc(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                                      int block_cols, int64_t ld_in, void* out, int64_t ld_out, int dtype);
/* Algorithmic HBM bytes and flops of one product of width block_cols. */
double amfls_operator_product_bytes(const amfls_operator* op, int block_cols, int transpose, int dtype);
double amfls_operator_product_flops(const amfls_operator* op, int block_cols);
/* Kernel variant per (width, direction); -1 restores the default.  CSR
 * operators: an SpMM variant index; choosing the width-one "slab" variant
 * builds the operator's slab-tiled copy of A (or A^T) first, synchronously,
 * and fails with AMFLS_ERR_LENGTH beyond its limits (more than 1024 slabs
 * of 8192 columns, 2^31 entries).  Dense row-major operators: 0 the
 * vector-ALU kernel (widths <= 8; the default for one column), 1 the matrix
 * cores (the default for wider blocks). */
int amfls_operator_set_variant(amfls_operator* op, int block_cols, int transpose, int variant);
/* SpMM autotuner (engine extension; the reference's CSR apply has a single
 * loop, src/algorithms/operators/sparse_csr_operator.cpp:203-269): for each
 * of `count` block widths, times the kernel variants whose column lanes fit
 * the width on this matrix (median of `reps` launches, blocks of dtype
 * AMFLS_F64 or AMFLS_F32) for A and A^T and keeps the fastest.  With width
 * one among the widths it first builds the slab-tiled copies where x is
 * reused enough within a panel (staging it moves at most 8 times the entry
 * bytes), so the slab kernel is a candidate there.  chosen (may be NULL)
 * receives 2 * count variant indices, (A, A^T) per width. */
int amfls_operator_autotune(amfls_context* ctx, amfls_operator* op, const int32_t* widths, int32_t count,
                            int dtype, int32_t reps, int32_t* chosen);
/* Per-product kernel timing (off by default): CSR products are timed by the
 * start and end timestamps of their kernel dispatches, other products by
 * HIP events around them on the engine stream.  kernel_stats returns and
 * resets: launches, total seconds and total algorithmic bytes. */
int amfls_operator_profile(amfls_operator* op, int enable);
int amfls_operator_kernel_stats(amfls_operator* op, int64_t* launches, double* seconds, double* bytes);
/* The same totals over one class of the products timed since the last
 * kernel_stats (it does not reset): direction (transpose 0: A X, 1: A^T Y),
 * block widths in [min_cols, max_cols] and block type (AMFLS_F64 or
 * AMFLS_F32).  A per-kernel view for rooflines: each class runs one product
 * kernel.  No counterpart in the reference (its operator seconds are host
 * wall time, matrix_operator.hpp). */
int amfls_operator_kernel_stats_class(amfls_operator* op, int transpose, int32_t min_cols, int32_t max_cols,
                                      int dtype, int64_t* launches, double* seconds, double* bytes);
int amfls_spmm_variant_count(void);
const char* amfls_spmm_variant_name(int variant);
int amfls_spmm_variant_supports(int variant, int block_cols);
/* The variant used when none is set: lanes sized to the block width and the
 * mean number of entries per row of the product's matrix (A or A^T). */
int amfls_spmm_default_variant(int block_cols, double mean_row_entries);

/* b: rows values, device pointer if b_on_device != 0 else host pointer. */
int amfls_solve(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                const amfls_options* options, amfls_result** out);
/* Regularization sweep (engine extension; the reference solves one lambda
 * per solve_amfls call, src/algorithms/amfls/amfls.cpp:57-661).  The block
 * Golub-Kahan basis of (A, b) does not depend on lambda: one process runs for
 * the smallest of lambdas[count] and, at each of its checkpoints, the basis
 * is also evaluated and certified (FP64) for every other value not yet
 * certified; a value still uncertified when the process stops is solved on
 * its own.  options->regularization is ignored; precision FP64 or FP32;
 * recording width policies are not accepted.  results[count] receive one
 * result per lambda (statistics of swept results are the shared process's). */
int amfls_solve_sweep(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                      const double* lambdas, int32_t count, const amfls_options* options, amfls_result** results);
/* Independent CSR problems solved concurrently (engine extension): `workers`
 * host threads, each with its own stream, take problems in order and run
 * each problem's regularization sweep over its lambdas_per_problem values
 * (lambdas is problems x lambdas_per_problem, row-major).  b_dev holds device
 * right-hand sides.  results (problems x lambdas_per_problem) as
 * amfls_solve_sweep.  An error in any problem fails the call. */
int amfls_solve_batch(amfls_context* ctx, int32_t problems, const amfls_operator* const* ops,
                      const double* const* b_dev, const double* lambdas, int32_t lambdas_per_problem,
                      const amfls_options* options, int32_t workers, amfls_result** results);
int amfls_solve_fixed_rbgk(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                           double regularization, double tolerance, double failure_probability,
                           int32_t auxiliary_width, int32_t maximum_depth, int32_t maximum_basis_size,
                           uint64_t seed, uint64_t stream, int32_t clock_mode, amfls_result** out);
int amfls_solve_lsqr(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                     double regularization, double tolerance, int32_t maximum_iterations, int32_t clock_mode,
                     amfls_result** out);

/* Certified validation of a given candidate x (length cols) for
 * min ||Ax - b||^2 + lambda ||x||^2: fresh FP64 products, outward-rounded
 * norm intervals and rounding radii, the backward-error / ridge-energy
 * bounds, the long-double refinement branch and the one-step ridge
 * correction; the result carries x, the certificate metrics and the
 * validation counters (no trace rows).  Replaces the reference's
 * validate_original_candidate
 * (src/algorithms/mathematics/krylov/candidate_validation.cpp:1176-1378),
 * which the reference runs on every candidate of a solve. */
int amfls_validate_candidate(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                             const double* x, int x_on_device, double regularization, double tolerance,
                             amfls_result** out);
int amfls_result_sizes(const amfls_result* result, int64_t* cols, int64_t* trace_rows);
/* Host buffers: solution[cols], metrics[21], counters[36], timings[9],
 * trace_float[trace_rows * 29], trace_int[trace_rows * 4] (row-major). */
int amfls_result_copy(const amfls_result* result, double* solution, double* metrics, int64_t* counters,
                      double* timings, double* trace_float, int64_t* trace_int, int64_t* status,
                      int64_t* stop_reason);
/* Engine extension per trace row: interval operator seconds, interval local
 * seconds, active width of the interval (trace_rows * 3). */
int amfls_result_trace_extension(const amfls_result* result, double* out);
int amfls_result_refinement_steps(const amfls_result* result, int32_t* steps);
/* Forward-error estimate of a mixed-precision result (NaN where none). */
int amfls_result_forward_error_estimate(const amfls_result* result, double* estimate);
/* 1 when a mixed-precision solve with a forward-error target missed the target
 * in its FP64 refinement and the returned result is the FP64 engine's (the
 * trace and product counters then cover both runs), else 0.  No counterpart
 * in the reference, which has no mixed-precision path. */
int amfls_result_fp64_fallback(const amfls_result* result, int32_t* fell_back);
int amfls_result_destroy(amfls_result* result);

/* Learned width controller (engine extension; it replaces the reference's
 * widening predicate, src/algorithms/amfls/amfls.cpp:322-358 and the cost
 * gate of amfls_internal.hpp:318-324, when options.width_policy is
 * AMFLS_WIDTH_LEARNED; the FP64 certificate still decides termination).
 * A 12 -> 64 -> 64 -> 1 ReLU MLP in FP32: params holds AMFLS_WIDTH_PARAMS
 * values (W1[64x12], b1[64], W2[64x64], b2[64], W3[64], b3), norm the input
 * means then inverse standard deviations (2 x 12).  A model uses the
 * context it was created on and must be destroyed before it.
 * AMFLS_WIDTH_RECORD solves with the heuristic and records, per widening
 * decision, the 12 model inputs and the heuristic's choice (1 = widen). */
int amfls_width_model_create(amfls_context* ctx, const float* params, const float* norm, amfls_width_model** out);
int amfls_width_model_destroy(amfls_width_model* model);
int amfls_width_model_get_params(amfls_width_model* model, float* params);
int amfls_width_model_set_params(amfls_width_model* model, const float* params);
/* Logits of n host feature rows (n x 12) into host logits[n]. */
int amfls_width_model_logits(amfls_width_model* model, const float* features, int64_t n, float* logits);
/* Mean binary cross-entropy of device features (n x 12) and labels (n) and
 * its gradient (host grad[AMFLS_WIDTH_PARAMS], may be NULL). */
int amfls_width_model_gradient(amfls_width_model* model, const float* features, const float* labels, int64_t n,
                               float* grad, float* loss);
/* One Adam step on a device batch; *loss is the batch loss before the step. */
int amfls_width_model_train_step(amfls_width_model* model, const float* features, const float* labels, int64_t n,
                                 float learning_rate, float* loss);
/* Per-input box (lo[12], hi[12]) of the data the model was trained on.  A
 * decision whose inputs leave the box is taken by the heuristic instead, so
 * the model never extrapolates (the default box is unbounded). */
int amfls_width_model_set_domain(amfls_width_model* model, const float* lo, const float* hi);
/* Decisions taken by the model and handed to the heuristic so far. */
int amfls_width_model_decisions(const amfls_width_model* model, int64_t* by_model, int64_t* by_heuristic);
/* The model used by AMFLS_WIDTH_LEARNED solves on this context (NULL clears). */
int amfls_context_set_width_model(amfls_context* ctx, amfls_width_model* model);
/* Width-one level batches of later solves on this context are captured into
 * a HIP graph (stream capture of the batch's products and level-chain
 * kernels) and launched as one graph (enabled != 0), or enqueued kernel by
 * kernel (0, the default).  No counterpart in the reference, which has no
 * device queue; results are identical either way.  Operators timed by
 * events (dense products on the wall clock) and operators being profiled
 * keep the direct enqueue. */
int amfls_context_set_level_graphs(amfls_context* ctx, int enabled);
/* Recorded decisions of an AMFLS_WIDTH_RECORD solve: rows x 13 floats
 * (out may be NULL to query rows). */
int amfls_result_width_records(const amfls_result* result, float* out, int64_t* rows);

/* Counter-addressed synthetic CSR on the device (engine extension for the
 * batched benchmark inputs; its counters follow the splitmix64 / Box-Muller
 * chain of reference random/gaussian.cpp:13-35).  generate_rows writes
 * row_ptr (rows + 1 device int64) for per-row lengths uniform in
 * [nnz_lo, min(nnz_hi, cols)] (nnz_hi <= 64) and returns *nnz (waits);
 * generate_entries then writes ascending distinct uniform column indices
 * (nnz device int32) and N(0,1) values times log-uniform row scales in
 * [scale_lo, scale_hi] (nnz device values of value_type) and waits. */
int amfls_csr_generate_rows(amfls_context* ctx, int64_t rows, int64_t cols, int32_t nnz_lo, int32_t nnz_hi,
                            uint64_t seed, int64_t* row_ptr, int64_t* nnz);
int amfls_csr_generate_entries(amfls_context* ctx, int64_t rows, int64_t cols, uint64_t seed, double scale_lo,
                               double scale_hi, const int64_t* row_ptr, int32_t* col_idx, void* values,
                               int value_type);

/* Counter-addressed Gaussian fill (reference random/gaussian.cpp:25-55)
 * of a row-major device block; dtype AMFLS_F64 or AMFLS_F32. */
int amfls_gaussian_fill(amfls_context* ctx, void* y, int64_t rows, int32_t cols, int64_t ld, int dtype,
                        uint64_t seed, uint64_t stream, uint64_t first_col);

/* The engine's FP64 block kernels on row-major device blocks (the dense
 * BLAS of the reference's Krylov workspace, linalg/blas_lapack.cpp dgemm
 * and dnrm2 call sites; exposed so callers and tests can check them):
 *   block_gemm_nn: Y (m x n, ldy) = alpha A C + beta Y, A m x k (lda),
 *                  C k x n (ldc);
 *   block_gemm_tn: C (k x n, ldc) = alpha A^T B, A m x k (lda), B m x n (ldb);
 *   block_rowmul:  Y[i, 0:r] = X[i, 0:w] M, M w x r (ld r); Y may be X;
 *   block_stats:   out[3] (host) = (max |v|, sum (|v| / 2^e)^2, nonfinite
 *                  count), e the frexp exponent of the max.
 * Each call waits for its kernels. */
int amfls_block_gemm_nn(amfls_context* ctx, int64_t m, int32_t n, int32_t k, const double* a, int64_t lda,
                        const double* c, int64_t ldc, double* y, int64_t ldy, double alpha, double beta);
int amfls_block_gemm_tn(amfls_context* ctx, int64_t m, int32_t k, int32_t n, const double* a, int64_t lda,
                        const double* b, int64_t ldb, double* c, int64_t ldc, double alpha);
int amfls_block_rowmul(amfls_context* ctx, int64_t rows, int32_t w, int32_t r, const double* x, int64_t ldx,
                       const double* m, double* y, int64_t ldy);
int amfls_block_stats(amfls_context* ctx, int64_t rows, int32_t cols, const double* x, int64_t ldx, double* out);
/* Rank reveal of one device block W (rows x w, row-major, FP64 or FP32) as
 * the search factors every new block: CholeskyQR passes, singular values
 * of the triangular factor, numerical rank at the cutoff (the reference's
 * rank_revealing_cutoff scaled by ||W||_F; fp32_rank_factor > 0 replaces
 * the FP32 rule's factor).  W is overwritten with its orthonormal factor.
 * Outputs: rank, cutoff, the norm of the discarded singular values, and the
 * w retained singular values in decreasing order (zeros past the rank).
 * Replaces the SVD step of the reference's orthogonalize_and_factor
 * (src/algorithms/mathematics/krylov/krylov_workspace.cpp:263-309); for
 * tests and diagnostics. */
int amfls_block_rank_reveal(amfls_context* ctx, int64_t rows, int32_t w, void* x, int64_t ldx, int dtype,
                            int32_t max_rank, double fp32_rank_factor, int32_t* rank, double* cutoff,
                            double* discarded, double* singular_values);

/* The same products on FP32 blocks (the mixed-precision search's local
 * work): A, B, X, Y FP32; C and M FP64; block_gemm_tn_f32 accumulates in
 * FP64 and writes FP64 C. */
int amfls_block_gemm_nn_f32(amfls_context* ctx, int64_t m, int32_t n, int32_t k, const float* a, int64_t lda,
                            const double* c, int64_t ldc, float* y, int64_t ldy, double alpha, double beta);
int amfls_block_gemm_tn_f32(amfls_context* ctx, int64_t m, int32_t k, int32_t n, const float* a, int64_t lda,
                            const float* b, int64_t ldb, double* c, int64_t ldc, double alpha);
int amfls_block_rowmul_f32(amfls_context* ctx, int64_t rows, int32_t w, int32_t r, const float* x, int64_t ldx,
                           const double* m, float* y, int64_t ldy);

#ifdef __cplusplus
}
#endif

#endif /* AMFLS_HIP_H */
