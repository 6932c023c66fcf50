// extern "C" boundary of the AMFLS HIP engine (declared in include/amfls_hip.h).
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../../include/amfls_hip.h"
#include "solver.hpp"
#include "validation.hpp"
#include "workspace.hpp"
#include "width_model.hpp"

using namespace amflsx;

struct amfls_width_model {
    std::unique_ptr<WidthModel> m;
};

struct amfls_context {
    Context ctx;
    amfls_width_model* width_model = nullptr;  // not owned
};

struct amfls_operator {
    std::unique_ptr<DeviceOperator> op;
};

struct amfls_result {
    LeastSquaresResult r;
    std::vector<float> width_records;  // rows x (AMFLS_WIDTH_INPUTS + 1) in record mode
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
    try {
        f();
        g_last_error.clear();
        return AMFLS_OK;
    } catch (const HipError& e) {
        g_last_error = e.what();
        return AMFLS_ERR_HIP;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return AMFLS_ERR_INVALID_ARGUMENT;
    } catch (const std::length_error& e) {
        g_last_error = e.what();
        return AMFLS_ERR_LENGTH;
    } catch (const std::overflow_error& e) {
        g_last_error = e.what();
        return AMFLS_ERR_LENGTH;
    } catch (const std::logic_error& e) {
        g_last_error = e.what();
        return AMFLS_ERR_LOGIC;
    } catch (const std::bad_alloc& e) {
        g_last_error = "host allocation failed";
        return AMFLS_ERR_RUNTIME;
    } catch (const std::runtime_error& e) {
        g_last_error = e.what();
        return AMFLS_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return AMFLS_ERR_UNKNOWN;
    } catch (...) {
        g_last_error = "unknown error";
        return AMFLS_ERR_UNKNOWN;
    }
}

void require(bool c, const char* msg) {
    if (!c) throw std::invalid_argument(msg);
}

ValueType to_vt(int t) {
    if (t == AMFLS_F64) return ValueType::f64;
    if (t == AMFLS_F32) return ValueType::f32;
    if (t == AMFLS_BF16) return ValueType::bf16;
    throw std::invalid_argument("unknown value type");
}

// Device copy of b (owned when b was given on the host).
struct RhsHolder {
    DevBuf<double> own;
    const double* ptr = nullptr;
    RhsHolder(Context& ctx, const double* b, int64_t m, bool on_device) {
        require(b != nullptr, "right-hand side cannot be null");
        if (on_device) {
            ptr = b;
        } else {
            own.reset(static_cast<size_t>(m));
            AMFLS_HIP_CHECK(hipMemcpyAsync(own.get(), b, m * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
            ctx.sync();
            ptr = own.get();
        }
    }
};

template <typename T>
void transpose_block(hipStream_t s, int64_t rows, int cols, const T* src_colmajor, T* dst_rowmajor) {
    // column-major (rows x cols, ld = rows) -> row-major (ld = cols):
    // element (i, j) of the column-major block is the row-major element
    // (j, i) of a cols x rows matrix, so copy_block over that view.
    // copy_block reads X[i * ldx + j]; use the transposed strides via a
    // per-column loop (cols is the block width, small).
    for (int j = 0; j < cols; ++j)
        copy_block<T, T>(s, rows, 1, src_colmajor + j * rows, 1, dst_rowmajor + j, cols, 1.0);
}

template <typename T>
void untranspose_block(hipStream_t s, int64_t rows, int cols, const T* src_rowmajor, T* dst_colmajor) {
    for (int j = 0; j < cols; ++j)
        copy_block<T, T>(s, rows, 1, src_rowmajor + j, cols, dst_colmajor + j * rows, 1, 1.0);
}

template <typename T>
void apply_colmajor(Context& ctx, const DeviceOperator& op, bool transpose, const T* in, int s, T* out) {
    const int64_t in_rows = transpose ? op.rows() : op.cols();
    const int64_t out_rows = transpose ? op.cols() : op.rows();
    DevBuf<T> a(static_cast<size_t>(in_rows * s)), b(static_cast<size_t>(out_rows * s));
    transpose_block<T>(ctx.stream, in_rows, s, in, a.get());
    if (transpose) op.apply_t(ctx.stream, a.get(), s, s, b.get(), s);
    else op.apply(ctx.stream, a.get(), s, s, b.get(), s);
    untranspose_block<T>(ctx.stream, out_rows, s, b.get(), out);
    ctx.sync();
}

}  // namespace

extern "C" {

const char* amfls_last_error(void) { return g_last_error.c_str(); }
int amfls_version(void) { return 10000; }

void amfls_options_default(amfls_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->regularization = 0.0;
    o->tolerance = 1e-8;
    o->failure_probability = 1e-6;
    o->precision = AMFLS_PRECISION_FP64;
    o->clock_mode = AMFLS_CLOCK_WALL;
    o->synthetic_operator_unit = 1e-3;
    o->synthetic_local_unit = 1e-6;
    o->maximum_refinements = 8;
    o->inner_tolerance = 0.0;
    o->width_policy = AMFLS_WIDTH_HEURISTIC;
    o->forward_tolerance = 0.0;
    o->maximum_refinement_iterations = 0;
    o->width_script = nullptr;
    o->width_script_length = 0;
}

int amfls_context_create(amfls_context** out) {
    return guard([&] {
        require(out != nullptr, "output pointer is null");
        *out = new amfls_context();
        // Set-up of the allocation-free product on this device: resolving
        // the address of its static array loads the array (once per device)
        // here, so the entry point itself never allocates.
        (void)noalloc_scratch();
    });
}

int amfls_context_destroy(amfls_context* ctx) {
    return guard([&] { delete ctx; });
}

int amfls_context_synchronize(amfls_context* ctx) {
    return guard([&] {
        require(ctx != nullptr, "context is null");
        ctx->ctx.sync();
    });
}

int amfls_csr_create(amfls_context* ctx, int64_t rows, int64_t cols, int64_t nnz, const int64_t* row_ptr,
                     const int32_t* col_idx, const void* values, int value_type, amfls_operator** out) {
    return guard([&] {
        require(ctx != nullptr && out != nullptr, "null context or output");
        auto h = std::make_unique<amfls_operator>();
        h->op = std::make_unique<CsrOperator>(rows, cols, nnz, row_ptr, col_idx, values, to_vt(value_type), ctx->ctx);
        *out = h.release();
    });
}

int amfls_dense_create(amfls_context* ctx, int64_t rows, int64_t cols, const void* values, int value_type,
                       int row_major, int64_t ld, amfls_operator** out) {
    return guard([&] {
        require(ctx != nullptr && out != nullptr, "null context or output");
        auto h = std::make_unique<amfls_operator>();
        h->op = std::make_unique<DenseOperator>(rows, cols, values, to_vt(value_type), row_major != 0, ld, ctx->ctx);
        *out = h.release();
    });
}

int amfls_host_operator_create(amfls_context* ctx, int64_t rows, int64_t cols, amfls_host_apply_fn apply,
                               amfls_host_apply_fn apply_transpose, void* user, double norm_upper,
                               double norm_lower, int64_t steps_apply, int64_t steps_transpose,
                               amfls_operator** out) {
    return guard([&] {
        require(ctx != nullptr && out != nullptr, "null context or output");
        ValidationModel m;
        m.upper = norm_upper;
        m.lower = norm_lower;
        m.steps_apply = steps_apply;
        m.steps_transpose = steps_transpose;
        auto h = std::make_unique<amfls_operator>();
        h->op = std::make_unique<HostOperator>(rows, cols, apply, apply_transpose, user, m);
        *out = h.release();
    });
}

int amfls_host_operator_create_ex(amfls_context* ctx, int64_t rows, int64_t cols, const amfls_host_callbacks* cb,
                                  double norm_upper, double norm_lower, int64_t steps_apply, int64_t steps_transpose,
                                  amfls_operator** out) {
    return guard([&] {
        require(ctx != nullptr && out != nullptr && cb != nullptr, "null context, callbacks or output");
        ValidationModel m;
        m.upper = norm_upper;
        m.lower = norm_lower;
        m.steps_apply = steps_apply;
        m.steps_transpose = steps_transpose;
        HostHooks hooks;
        hooks.cost = cb->relative_block_product_cost;
        hooks.refine = cb->apply_refinement;
        hooks.refine_t = cb->apply_transpose_refinement;
        auto h = std::make_unique<amfls_operator>();
        h->op = std::make_unique<HostOperator>(rows, cols, cb->apply, cb->apply_transpose, cb->user, m, hooks);
        *out = h.release();
    });
}

int amfls_operator_destroy(amfls_operator* op) {
    return guard([&] { delete op; });
}

int amfls_operator_shape(const amfls_operator* op, int64_t* rows, int64_t* cols) {
    return guard([&] {
        require(op && rows && cols, "null argument");
        *rows = op->op->rows();
        *cols = op->op->cols();
    });
}

int amfls_operator_validation_model(const amfls_operator* op, double* upper, double* lower, int64_t* sa,
                                    int64_t* st) {
    return guard([&] {
        require(op && upper && lower && sa && st, "null argument");
        const ValidationModel m = op->op->validation_model();
        *upper = m.upper;
        *lower = m.lower;
        *sa = m.steps_apply;
        *st = m.steps_transpose;
    });
}

double amfls_operator_relative_block_product_cost(const amfls_operator* op, int s) {
    if (!op) return kInfD;
    return op->op->relative_block_product_cost(s);
}

int amfls_operator_apply(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                         int block_cols, void* out, int dtype) {
    return guard([&] {
        require(ctx && op && in && out, "null argument");
        require(block_cols > 0, "block_cols must be positive");
        if (dtype == AMFLS_F64)
            apply_colmajor<double>(ctx->ctx, *op->op, transpose != 0, static_cast<const double*>(in), block_cols,
                                   static_cast<double*>(out));
        else if (dtype == AMFLS_F32)
            apply_colmajor<float>(ctx->ctx, *op->op, transpose != 0, static_cast<const float*>(in), block_cols,
                                  static_cast<float*>(out));
        else
            throw std::invalid_argument("block dtype must be float64 or float32");
    });
}

int amfls_operator_apply_rows(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                              int block_cols, int64_t ld_in, void* out, int64_t ld_out, int dtype) {
    return guard([&] {
        require(ctx && op && in && out, "null argument");
        require(block_cols > 0 && ld_in >= block_cols && ld_out >= block_cols, "invalid block shape");
        hipStream_t s = ctx->ctx.stream;
        if (dtype == AMFLS_F64) {
            if (transpose) op->op->apply_t(s, static_cast<const double*>(in), block_cols, ld_in, static_cast<double*>(out), ld_out);
            else op->op->apply(s, static_cast<const double*>(in), block_cols, ld_in, static_cast<double*>(out), ld_out);
        } else if (dtype == AMFLS_F32) {
            if (transpose) op->op->apply_t(s, static_cast<const float*>(in), block_cols, ld_in, static_cast<float*>(out), ld_out);
            else op->op->apply(s, static_cast<const float*>(in), block_cols, ld_in, static_cast<float*>(out), ld_out);
        } else {
            throw std::invalid_argument("block dtype must be float64 or float32");
        }
    });
}

int amfls_operator_apply_rows_noalloc(amfls_context* ctx, const amfls_operator* op, int transpose, const void* in,
                                      int block_cols, int64_t ld_in, void* out, int64_t ld_out, int dtype) {
    return guard([&] {
        require(ctx && op && in && out, "null argument");
        require(block_cols > 0 && ld_in >= block_cols && ld_out >= block_cols, "invalid block shape");
        require(dtype == AMFLS_F64 || dtype == AMFLS_F32, "block dtype must be float64 or float32");
        Scratch fixed = noalloc_scratch();
        if (!op->op->product_noalloc(ctx->ctx.stream, transpose != 0, in, block_cols, ld_in, out, ld_out,
                                     dtype == AMFLS_F64, fixed))
            throw std::invalid_argument("the allocation-free product supports device CSR and dense operators");
    });
}

double amfls_operator_product_bytes(const amfls_operator* op, int s, int transpose, int dtype) {
    if (!op) return 0.0;
    return op->op->product_bytes(s, transpose != 0, dtype == AMFLS_F64 ? 8 : 4);
}

double amfls_operator_product_flops(const amfls_operator* op, int s) {
    if (!op) return 0.0;
    return op->op->product_flops(s);
}

int amfls_operator_set_variant(amfls_operator* op, int block_cols, int transpose, int variant) {
    return guard([&] {
        require(op != nullptr, "operator is null");
        require(variant < csr_spmm_variant_count(), "unknown variant");
        op->op->set_variant(block_cols, transpose != 0, variant);
    });
}

int amfls_operator_profile(amfls_operator* op, int enable) {
    return guard([&] {
        require(op != nullptr, "operator is null");
        op->op->set_profiling(enable != 0);
    });
}

int amfls_operator_kernel_stats(amfls_operator* op, int64_t* launches, double* seconds, double* bytes) {
    return guard([&] {
        require(op && launches && seconds && bytes, "null argument");
        long long l = 0;
        op->op->kernel_stats(&l, seconds, bytes);
        *launches = l;
    });
}

int amfls_operator_kernel_stats_class(amfls_operator* op, int transpose, int32_t min_cols, int32_t max_cols,
                                      int dtype, int64_t* launches, double* seconds, double* bytes) {
    return guard([&] {
        require(op && launches && seconds && bytes, "null argument");
        require(dtype == AMFLS_F64 || dtype == AMFLS_F32, "block products are float64 or float32");
        require(min_cols >= 1 && min_cols <= max_cols, "width range is empty");
        long long l = 0;
        op->op->kernel_stats_class(transpose != 0, min_cols, max_cols, dtype == AMFLS_F64 ? 8 : 4, &l, seconds,
                                   bytes);
        *launches = l;
    });
}

int amfls_operator_autotune(amfls_context* ctx, amfls_operator* op, const int32_t* widths, int32_t count,
                            int dtype, int32_t reps, int32_t* chosen) {
    return guard([&] {
        require(ctx && op && widths && count > 0, "invalid argument");
        require(dtype == AMFLS_F64 || dtype == AMFLS_F32, "autotune blocks are float64 or float32");
        auto* csr = dynamic_cast<CsrOperator*>(op->op.get());
        require(csr != nullptr, "autotuning applies to CSR operators");
        const std::vector<int> v = csr->autotune(ctx->ctx, std::vector<int>(widths, widths + count),
                                                 dtype == AMFLS_F64, reps);
        if (chosen)
            for (size_t i = 0; i < v.size(); ++i) chosen[i] = v[i];
    });
}

int amfls_spmm_variant_count(void) { return csr_spmm_variant_count(); }
const char* amfls_spmm_variant_name(int v) { return csr_spmm_variant_name(v); }
int amfls_spmm_variant_supports(int v, int s) { return csr_spmm_variant_supports(v, s) ? 1 : 0; }
int amfls_spmm_default_variant(int s, double mean_row_entries) { return csr_spmm_default_variant(s, mean_row_entries); }

}  // extern "C"

namespace {

SolveOptions solve_options(const amfls_options* options) {
    SolveOptions o;
    o.regularization = options->regularization;
    o.tolerance = options->tolerance;
    o.failure_probability = options->failure_probability;
    o.maximum_epochs = options->maximum_epochs;
    o.maximum_basis_size = options->maximum_basis_size;
    o.seed = options->seed;
    o.stream = options->stream;
    o.maximum_depth = options->maximum_depth;
    o.maximum_auxiliary_width = options->maximum_auxiliary_width;
    require(options->precision >= AMFLS_PRECISION_FP64 && options->precision <= AMFLS_PRECISION_FP32,
            "unknown precision");
    o.precision = static_cast<Precision>(options->precision);
    o.clock_mode = options->clock_mode == AMFLS_CLOCK_SYNTHETIC ? ClockMode::synthetic : ClockMode::wall;
    o.synthetic_operator_unit = options->synthetic_operator_unit;
    o.synthetic_local_unit = options->synthetic_local_unit;
    o.maximum_refinements = options->maximum_refinements;
    o.inner_tolerance = options->inner_tolerance;
    o.width_policy = options->width_policy;
    require(std::isfinite(options->forward_tolerance) && options->forward_tolerance >= 0.0 &&
                options->forward_tolerance < 1.0,
            "forward_tolerance must lie in [0, 1)");
    require(options->maximum_refinement_iterations >= 0, "maximum_refinement_iterations must be nonnegative");
    o.forward_tolerance = options->forward_tolerance;
    o.maximum_refinement_iterations = options->maximum_refinement_iterations;
    require(o.width_policy >= AMFLS_WIDTH_HEURISTIC && o.width_policy <= AMFLS_WIDTH_RECORD_WIDEN,
            "unknown width policy");
    return o;
}

// Widening predicate of a width policy: the reference heuristic (empty),
// the learned model, constant choices, or the heuristic / always-widen with
// every decision's model inputs recorded into `records`.
// The policy of `base` with its first decisions replaced by a script (0
// hold, 1 widen, other: the base policy).  The base policy still sees every
// decision, so a recording base records the scripted ones too.
WidthPolicy scripted_policy(WidthPolicy base, const int8_t* script, int32_t length) {
    require(length >= 0 && (length == 0 || script != nullptr), "invalid width script");
    if (length == 0) return base;
    auto steps = std::make_shared<std::vector<int8_t>>(script, script + length);
    auto next = std::make_shared<size_t>(0);
    return [base, steps, next](const WidthFeatures& f) {
        const size_t i = (*next)++;
        const bool chosen = base ? base(f) : f.heuristic;
        if (i < steps->size() && ((*steps)[i] == 0 || (*steps)[i] == 1))
            return (*steps)[i] == 1 && f.active_width + f.proposed_increment <= 64;
        return chosen;
    };
}

WidthPolicy width_policy(const amfls_context* ctx, int code, std::vector<float>* records) {
    if (code == AMFLS_WIDTH_LEARNED) {
        require(ctx->width_model != nullptr, "the learned width policy needs a model set on the context");
        WidthModel* m = ctx->width_model->m.get();
        // the model decides within the widths it was trained on (blocks of at most 64 columns)
        return [m](const WidthFeatures& f) { return f.active_width + f.proposed_increment <= 64 && m->widen(f); };
    }
    if (code == AMFLS_WIDTH_WIDEN)  // widen at every opportunity up to the widest benchmarked block (64)
        return [](const WidthFeatures& f) { return f.active_width + f.proposed_increment <= 64; };
    if (code == AMFLS_WIDTH_HOLD) return [](const WidthFeatures&) { return false; };
    if (code == AMFLS_WIDTH_RECORD || code == AMFLS_WIDTH_RECORD_WIDEN) {
        require(records != nullptr, "recording width policies apply to single solves");
        const bool widen_always = code == AMFLS_WIDTH_RECORD_WIDEN;
        return [records, widen_always](const WidthFeatures& f) {
            const auto x = width_features(f);
            records->insert(records->end(), x.begin(), x.end());
            records->push_back(f.heuristic ? 1.0f : 0.0f);
            return widen_always ? f.active_width + f.proposed_increment <= 64 : f.heuristic;
        };
    }
    return WidthPolicy{};
}

}  // namespace

extern "C" {

int amfls_solve(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                const amfls_options* options, amfls_result** out) {
    return guard([&] {
        require(ctx && op && options && out, "null argument");
        const SolveOptions o = solve_options(options);
        RhsHolder rhs(ctx->ctx, b, op->op->rows(), b_on_device != 0);
        auto res = std::make_unique<amfls_result>();
        const WidthPolicy policy = scripted_policy(width_policy(ctx, o.width_policy, &res->width_records),
                                                   options->width_script, options->width_script_length);
        res->r = solve_amfls(*op->op, rhs.ptr, o, ctx->ctx, policy ? &policy : nullptr);
        *out = res.release();
    });
}

int amfls_solve_sweep(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                      const double* lambdas, int32_t count, const amfls_options* options, amfls_result** results) {
    return guard([&] {
        require(ctx && op && options && results && lambdas && count > 0, "invalid argument");
        const SolveOptions o = solve_options(options);
        RhsHolder rhs(ctx->ctx, b, op->op->rows(), b_on_device != 0);
        const WidthPolicy policy = width_policy(ctx, o.width_policy, nullptr);
        std::vector<LeastSquaresResult> rs = solve_amfls_sweep(
            *op->op, rhs.ptr, o, std::vector<double>(lambdas, lambdas + count), ctx->ctx, policy ? &policy : nullptr);
        std::vector<std::unique_ptr<amfls_result>> hold(static_cast<size_t>(count));
        for (int32_t j = 0; j < count; ++j) {
            hold[j] = std::make_unique<amfls_result>();
            hold[j]->r = std::move(rs[j]);
        }
        for (int32_t j = 0; j < count; ++j) results[j] = hold[j].release();
    });
}

int amfls_solve_batch(amfls_context* ctx, int32_t problems, const amfls_operator* const* ops,
                      const double* const* b_dev, const double* lambdas, int32_t lambdas_per_problem,
                      const amfls_options* options, int32_t workers, amfls_result** results) {
    return guard([&] {
        require(ctx && options && results && problems >= 0 && lambdas_per_problem > 0 && lambdas, "invalid argument");
        require(problems == 0 || (ops && b_dev), "null argument");
        const SolveOptions o = solve_options(options);
        const WidthPolicy policy = width_policy(ctx, o.width_policy, nullptr);
        std::vector<BatchProblem> batch(static_cast<size_t>(problems));
        for (int32_t i = 0; i < problems; ++i) {
            require(ops[i] != nullptr && b_dev[i] != nullptr, "null operator or right-hand side in the batch");
            require(dynamic_cast<const CsrOperator*>(ops[i]->op.get()) != nullptr,
                    "batched solves take CSR operators");
            batch[i].op = ops[i]->op.get();
            batch[i].b_dev = b_dev[i];
            batch[i].lambdas.assign(lambdas + static_cast<size_t>(i) * lambdas_per_problem,
                                    lambdas + static_cast<size_t>(i + 1) * lambdas_per_problem);
        }
        ctx->ctx.sync();  // inputs written on the caller's stream are complete
        auto rs = solve_amfls_batch(batch, o, workers, policy ? &policy : nullptr);
        const size_t total = static_cast<size_t>(problems) * lambdas_per_problem;
        std::vector<std::unique_ptr<amfls_result>> hold(total);
        for (size_t k = 0; k < total; ++k) {
            hold[k] = std::make_unique<amfls_result>();
            hold[k]->r = std::move(rs[k / lambdas_per_problem][k % lambdas_per_problem]);
        }
        for (size_t k = 0; k < total; ++k) results[k] = hold[k].release();
    });
}

int amfls_solve_fixed_rbgk(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                           double regularization, double tolerance, double failure_probability,
                           int32_t auxiliary_width, int32_t maximum_depth, int32_t maximum_basis_size,
                           uint64_t seed, uint64_t stream, int32_t clock_mode, amfls_result** out) {
    return guard([&] {
        require(ctx && op && out, "null argument");
        FixedRbgkOptions o;
        o.regularization = regularization;
        o.tolerance = tolerance;
        o.failure_probability = failure_probability;
        o.auxiliary_width = auxiliary_width;
        o.maximum_depth = maximum_depth;
        o.maximum_basis_size = maximum_basis_size;
        o.seed = seed;
        o.stream = stream;
        o.clock_mode = clock_mode == AMFLS_CLOCK_SYNTHETIC ? ClockMode::synthetic : ClockMode::wall;
        RhsHolder rhs(ctx->ctx, b, op->op->rows(), b_on_device != 0);
        auto res = std::make_unique<amfls_result>();
        res->r = solve_fixed_rbgk(*op->op, rhs.ptr, o, ctx->ctx);
        *out = res.release();
    });
}

int amfls_solve_lsqr(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                     double regularization, double tolerance, int32_t maximum_iterations, int32_t clock_mode,
                     amfls_result** out) {
    return guard([&] {
        require(ctx && op && out, "null argument");
        LsqrOptions o;
        o.regularization = regularization;
        o.tolerance = tolerance;
        o.maximum_iterations = maximum_iterations;
        o.clock_mode = clock_mode == AMFLS_CLOCK_SYNTHETIC ? ClockMode::synthetic : ClockMode::wall;
        RhsHolder rhs(ctx->ctx, b, op->op->rows(), b_on_device != 0);
        auto res = std::make_unique<amfls_result>();
        res->r = solve_lsqr(*op->op, rhs.ptr, o, ctx->ctx);
        *out = res.release();
    });
}

int amfls_validate_candidate(amfls_context* ctx, const amfls_operator* op, const double* b, int b_on_device,
                             const double* x, int x_on_device, double regularization, double tolerance,
                             amfls_result** out) {
    return guard([&] {
        require(ctx && op && out, "null argument");
        require(std::isfinite(regularization) && regularization >= 0.0,
                "regularization must be finite and nonnegative");
        require(std::isfinite(tolerance) && tolerance > 0.0, "tolerance must be finite and positive");
        const int64_t n = op->op->cols();
        RhsHolder rhs(ctx->ctx, b, op->op->rows(), b_on_device != 0);
        RhsHolder cand(ctx->ctx, x, n, x_on_device != 0);
        auto res = std::make_unique<amfls_result>();
        RunStatistics stats;
        ValidationEngine ve(*op->op, rhs.ptr, ctx->ctx, stats);
        res->r = ve.validate(cand.ptr, regularization, tolerance);
        res->r.solution.resize(static_cast<size_t>(n));
        AMFLS_HIP_CHECK(hipMemcpyAsync(res->r.solution.data(), cand.ptr, n * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->ctx.stream));
        ctx->ctx.sync();
        res->r.statistics = stats;
        *out = res.release();
    });
}

int amfls_result_sizes(const amfls_result* r, int64_t* cols, int64_t* rows) {
    return guard([&] {
        require(r && cols && rows, "null argument");
        *cols = static_cast<int64_t>(r->r.solution.size());
        *rows = static_cast<int64_t>(r->r.trace.size());
    });
}

int amfls_result_copy(const amfls_result* res, double* solution, double* metrics, int64_t* counters,
                      double* timings, double* trace_float, int64_t* trace_int, int64_t* status,
                      int64_t* stop_reason) {
    return guard([&] {
        require(res != nullptr, "result is null");
        const LeastSquaresResult& r = res->r;
        const RunStatistics& s = r.statistics;
        if (solution) std::memcpy(solution, r.solution.data(), r.solution.size() * sizeof(double));
        if (metrics) {
            const double m[AMFLS_METRICS_WIDTH] = {
                r.objective, r.residual_norm, r.augmented_residual_norm, r.normal_residual_norm, r.solution_norm,
                r.solution_energy_norm, r.operator_norm_lower_bound, r.augmented_operator_norm_lower_bound,
                r.compatible_backward_error_upper_bound, r.least_squares_backward_error_upper_bound,
                r.backward_error_upper_bound, r.relative_normal_residual_upper_bound, r.energy_error_upper_bound,
                r.relative_energy_error_upper_bound, r.objective_gap_upper_bound,
                r.ridge_base_energy_error_upper_bound, r.ridge_corrected_energy_error_upper_bound,
                r.ridge_correction_gamma, r.ridge_correction_z_h_z, r.ridge_correction_two_abs_z_t_q,
                r.ridge_correction_q_norm_squared};
            std::memcpy(metrics, m, sizeof(m));
        }
        if (counters) {
            const int64_t c[AMFLS_COUNTERS_WIDTH] = {
                s.a_columns, s.at_columns, s.a_block_calls, s.at_block_calls, s.search_a_columns,
                s.search_at_columns, s.validation_a_columns, s.validation_at_columns, r.iterations, r.depth,
                r.auxiliary_width, r.basis_rank, s.sketch_a_columns, s.sketch_at_columns, s.sketch_a_block_calls,
                s.sketch_at_block_calls, s.iterative_a_columns, s.iterative_at_columns, s.iterative_a_block_calls,
                s.iterative_at_block_calls, s.gaussian_random_columns, s.gaussian_random_values,
                s.search_a_block_calls, s.search_at_block_calls, s.validation_a_block_calls,
                s.validation_at_block_calls, s.base_validation_a_columns, s.base_validation_at_columns,
                s.base_validation_a_block_calls, s.base_validation_at_block_calls, s.ridge_correction_a_columns,
                s.ridge_correction_at_columns, s.ridge_correction_a_block_calls, s.ridge_correction_at_block_calls,
                s.gaussian_random_block_requests, static_cast<int64_t>(r.ridge_correction_disposition)};
            std::memcpy(counters, c, sizeof(c));
        }
        if (timings) {
            const double other = std::max(0.0, s.total_seconds - s.a_seconds - s.at_seconds -
                                                   s.orthogonalization_seconds - s.projected_solve_seconds);
            const double t[AMFLS_TIMINGS_WIDTH] = {s.total_seconds, s.a_seconds, s.at_seconds,
                                                   s.orthogonalization_seconds, s.projected_solve_seconds,
                                                   s.validation_seconds, other, s.base_validation_seconds,
                                                   s.ridge_correction_validation_seconds};
            std::memcpy(timings, t, sizeof(t));
        }
        for (size_t i = 0; i < r.trace.size(); ++i) {
            const IterationRecord& q = r.trace[i];
            if (trace_float) {
                const double f[AMFLS_TRACE_FLOAT_WIDTH] = {
                    static_cast<double>(q.epoch), static_cast<double>(q.depth), static_cast<double>(q.auxiliary_width),
                    static_cast<double>(q.basis_rank), static_cast<double>(q.a_columns),
                    static_cast<double>(q.at_columns), static_cast<double>(q.a_block_calls),
                    static_cast<double>(q.at_block_calls), q.objective, q.residual_norm, q.augmented_residual_norm,
                    q.normal_residual_norm, q.operator_norm_lower_bound, q.compatible_backward_error_upper_bound,
                    q.least_squares_backward_error_upper_bound, q.backward_error_upper_bound,
                    q.relative_normal_residual_upper_bound, q.relative_energy_error_upper_bound, q.solution_norm,
                    q.solution_energy_norm, q.augmented_operator_norm_lower_bound, q.energy_error_upper_bound,
                    q.objective_gap_upper_bound, q.ridge_base_energy_error_upper_bound,
                    q.ridge_corrected_energy_error_upper_bound, q.ridge_correction_gamma, q.ridge_correction_z_h_z,
                    q.ridge_correction_two_abs_z_t_q, q.ridge_correction_q_norm_squared};
                std::memcpy(trace_float + i * AMFLS_TRACE_FLOAT_WIDTH, f, sizeof(f));
            }
            if (trace_int) {
                const int64_t t[AMFLS_TRACE_INT_WIDTH] = {
                    static_cast<int64_t>(q.contract_status), static_cast<int64_t>(q.contract_stop_reason),
                    q.contract_passed ? 1 : 0, static_cast<int64_t>(q.ridge_correction_disposition)};
                std::memcpy(trace_int + i * AMFLS_TRACE_INT_WIDTH, t, sizeof(t));
            }
        }
        if (status) *status = static_cast<int64_t>(r.status);
        if (stop_reason) *stop_reason = static_cast<int64_t>(r.stop_reason);
    });
}

int amfls_result_trace_extension(const amfls_result* res, double* out) {
    return guard([&] {
        require(res && out, "null argument");
        for (size_t i = 0; i < res->r.trace.size(); ++i) {
            out[3 * i] = res->r.trace[i].interval_operator_seconds;
            out[3 * i + 1] = res->r.trace[i].interval_local_seconds;
            out[3 * i + 2] = res->r.trace[i].active_width;
        }
    });
}

int amfls_result_refinement_steps(const amfls_result* res, int32_t* steps) {
    return guard([&] {
        require(res && steps, "null argument");
        *steps = res->r.refinement_steps;
    });
}

int amfls_result_fp64_fallback(const amfls_result* res, int32_t* fell_back) {
    return guard([&] {
        require(res && fell_back, "null argument");
        *fell_back = res->r.fp64_fallback ? 1 : 0;
    });
}

int amfls_result_forward_error_estimate(const amfls_result* res, double* estimate) {
    return guard([&] {
        require(res && estimate, "null argument");
        *estimate = res->r.forward_error_estimate;
    });
}

int amfls_result_width_records(const amfls_result* res, float* out, int64_t* rows) {
    return guard([&] {
        require(res && rows, "null argument");
        const size_t w = AMFLS_WIDTH_INPUTS + 1;
        *rows = static_cast<int64_t>(res->width_records.size() / w);
        if (out) std::memcpy(out, res->width_records.data(), res->width_records.size() * sizeof(float));
    });
}

int amfls_width_model_create(amfls_context* ctx, const float* params, const float* norm, amfls_width_model** out) {
    return guard([&] {
        require(ctx && params && norm && out, "null argument");
        auto m = std::make_unique<amfls_width_model>();
        m->m = std::make_unique<WidthModel>(params, norm, ctx->ctx);
        *out = m.release();
    });
}

int amfls_width_model_destroy(amfls_width_model* model) {
    return guard([&] { delete model; });
}

int amfls_width_model_get_params(amfls_width_model* model, float* params) {
    return guard([&] {
        require(model && params, "null argument");
        model->m->params(params);
    });
}

int amfls_width_model_set_params(amfls_width_model* model, const float* params) {
    return guard([&] {
        require(model && params, "null argument");
        model->m->set_params(params);
    });
}

int amfls_width_model_logits(amfls_width_model* model, const float* features, int64_t n, float* logits) {
    return guard([&] {
        require(model && (n == 0 || (features && logits)) && n >= 0, "invalid argument");
        model->m->logits(features, n, logits);
    });
}

int amfls_width_model_gradient(amfls_width_model* model, const float* features, const float* labels, int64_t n,
                               float* grad, float* loss) {
    return guard([&] {
        require(model && features && labels && n > 0 && loss, "invalid argument");
        *loss = model->m->gradient(features, labels, n, grad);
    });
}

int amfls_width_model_train_step(amfls_width_model* model, const float* features, const float* labels, int64_t n,
                                 float learning_rate, float* loss) {
    return guard([&] {
        require(model && features && labels && n > 0 && loss, "invalid argument");
        require(std::isfinite(learning_rate) && learning_rate > 0.0f, "learning rate must be positive");
        *loss = model->m->train_step(features, labels, n, learning_rate);
    });
}

int amfls_width_model_set_domain(amfls_width_model* model, const float* lo, const float* hi) {
    return guard([&] {
        require(model && lo && hi, "null argument");
        model->m->set_domain(lo, hi);
    });
}

int amfls_width_model_decisions(const amfls_width_model* model, int64_t* by_model, int64_t* by_heuristic) {
    return guard([&] {
        require(model && by_model && by_heuristic, "null argument");
        *by_model = model->m->model_decisions();
        *by_heuristic = model->m->fallback_decisions();
    });
}

int amfls_context_set_width_model(amfls_context* ctx, amfls_width_model* model) {
    return guard([&] {
        require(ctx != nullptr, "null argument");
        ctx->width_model = model;
    });
}

int amfls_context_set_level_graphs(amfls_context* ctx, int enabled) {
    return guard([&] {
        require(ctx != nullptr, "null argument");
        ctx->ctx.level_graphs = enabled != 0;
    });
}

int amfls_result_destroy(amfls_result* res) {
    return guard([&] { delete res; });
}

int amfls_csr_generate_rows(amfls_context* ctx, int64_t rows, int64_t cols, int32_t nnz_lo, int32_t nnz_hi,
                            uint64_t seed, int64_t* row_ptr, int64_t* nnz) {
    return guard([&] {
        require(ctx && row_ptr && nnz, "null argument");
        *nnz = csr_generate_rows(ctx->ctx.stream, rows, cols, nnz_lo, nnz_hi, seed, row_ptr, ctx->ctx.scratch);
    });
}

int amfls_csr_generate_entries(amfls_context* ctx, int64_t rows, int64_t cols, uint64_t seed, double scale_lo,
                               double scale_hi, const int64_t* row_ptr, int32_t* col_idx, void* values,
                               int value_type) {
    return guard([&] {
        require(ctx && row_ptr && col_idx && values, "null argument");
        require(value_type == AMFLS_F64 || value_type == AMFLS_F32, "generated values are float64 or float32");
        csr_generate_entries(ctx->ctx.stream, rows, cols, seed, scale_lo, scale_hi, row_ptr, col_idx, values,
                             value_type == AMFLS_F64 ? ValueType::f64 : ValueType::f32);
        ctx->ctx.sync();
    });
}

int amfls_gaussian_fill(amfls_context* ctx, void* y, int64_t rows, int32_t cols, int64_t ld, int dtype,
                        uint64_t seed, uint64_t stream, uint64_t first_col) {
    return guard([&] {
        require(ctx && y, "null argument");
        require(rows >= 0 && cols >= 0 && ld >= cols, "invalid block shape");
        if (dtype == AMFLS_F64) gaussian_block<double>(ctx->ctx.stream, rows, cols, seed, stream, first_col, static_cast<double*>(y), ld);
        else if (dtype == AMFLS_F32) gaussian_block<float>(ctx->ctx.stream, rows, cols, seed, stream, first_col, static_cast<float*>(y), ld);
        else throw std::invalid_argument("dtype must be float64 or float32");
        ctx->ctx.sync();
    });
}

int amfls_block_gemm_nn(amfls_context* ctx, int64_t m, int32_t n, int32_t k, const double* a, int64_t lda,
                        const double* c, int64_t ldc, double* y, int64_t ldy, double alpha, double beta) {
    return guard([&] {
        require(ctx && (m == 0 || (a && y)) && (k == 0 || c), "null argument");
        require(m >= 0 && n >= 0 && k >= 0 && lda >= k && ldc >= n && ldy >= n, "invalid block shape");
        const Strided A{a, ValueType::f64, lda, 1};
        gemm_nn<double, double>(ctx->ctx.stream, m, n, k, A, c, ldc, y, ldy, alpha, beta);
        ctx->ctx.sync();
    });
}

int amfls_block_gemm_tn(amfls_context* ctx, int64_t m, int32_t k, int32_t n, const double* a, int64_t lda,
                        const double* b, int64_t ldb, double* c, int64_t ldc, double alpha) {
    return guard([&] {
        require(ctx && c && (m == 0 || (a && b)), "null argument");
        require(m >= 0 && n >= 0 && k >= 0 && lda >= k && ldb >= n && ldc >= n, "invalid block shape");
        const Strided A{a, ValueType::f64, lda, 1};
        gemm_tn<double>(ctx->ctx.stream, m, k, n, A, b, ldb, c, ValueType::f64, ldc, alpha, ctx->ctx.scratch);
        ctx->ctx.sync();
    });
}

int amfls_block_gemm_nn_f32(amfls_context* ctx, int64_t m, int32_t n, int32_t k, const float* a, int64_t lda,
                            const double* c, int64_t ldc, float* y, int64_t ldy, double alpha, double beta) {
    return guard([&] {
        require(ctx && (m == 0 || (a && y)) && (k == 0 || c), "null argument");
        require(m >= 0 && n >= 0 && k >= 0 && lda >= k && ldc >= n && ldy >= n, "invalid block shape");
        const Strided A{a, ValueType::f32, lda, 1};
        gemm_nn<float, double>(ctx->ctx.stream, m, n, k, A, c, ldc, y, ldy, alpha, beta);
        ctx->ctx.sync();
    });
}

int amfls_block_gemm_tn_f32(amfls_context* ctx, int64_t m, int32_t k, int32_t n, const float* a, int64_t lda,
                            const float* b, int64_t ldb, double* c, int64_t ldc, double alpha) {
    return guard([&] {
        require(ctx && c && (m == 0 || (a && b)), "null argument");
        require(m >= 0 && n >= 0 && k >= 0 && lda >= k && ldb >= n && ldc >= n, "invalid block shape");
        const Strided A{a, ValueType::f32, lda, 1};
        gemm_tn<float>(ctx->ctx.stream, m, k, n, A, b, ldb, c, ValueType::f64, ldc, alpha, ctx->ctx.scratch);
        ctx->ctx.sync();
    });
}

int amfls_block_rowmul_f32(amfls_context* ctx, int64_t rows, int32_t w, int32_t r, const float* x, int64_t ldx,
                           const double* m, float* y, int64_t ldy) {
    return guard([&] {
        require(ctx && m && (rows == 0 || (x && y)), "null argument");
        require(rows >= 0 && w > 0 && r > 0 && ldx >= w && ldy >= r, "invalid block shape");
        rowmul<float>(ctx->ctx.stream, rows, w, r, x, ldx, m, y, ldy);
        ctx->ctx.sync();
    });
}

int amfls_block_rowmul(amfls_context* ctx, int64_t rows, int32_t w, int32_t r, const double* x, int64_t ldx,
                       const double* m, double* y, int64_t ldy) {
    return guard([&] {
        require(ctx && m && (rows == 0 || (x && y)), "null argument");
        require(rows >= 0 && w >= 1 && r >= 1 && ldx >= w && ldy >= r, "invalid block shape");
        rowmul<double>(ctx->ctx.stream, rows, w, r, x, ldx, m, y, ldy);
        ctx->ctx.sync();
    });
}

int amfls_block_rank_reveal(amfls_context* ctx, int64_t rows, int32_t w, void* x, int64_t ldx, int dtype,
                            int32_t max_rank, double fp32_rank_factor, int32_t* rank, double* cutoff,
                            double* discarded, double* singular_values) {
    return guard([&] {
        require(ctx && x && rank && cutoff && discarded && singular_values, "null argument");
        require(rows >= 1 && w >= 1 && w <= 256 && ldx >= w && max_rank >= 0, "invalid block shape");
        require(dtype == AMFLS_F64 || dtype == AMFLS_F32, "rank reveal takes FP64 or FP32 blocks");
        // the operator only supplies the block's shape to the workspace
        DenseOperator op(rows, w, x, to_vt(dtype), true, ldx, ctx->ctx);
        DevBuf<double> b(static_cast<size_t>(rows));
        fill<double>(ctx->ctx.stream, b.get(), rows, 0.0);
        RunStatistics stats;
        ValidationEngine ve(op, b.get(), ctx->ctx, stats);
        std::vector<double> sv;
        auto run = [&](auto* block) {
            using T = std::remove_pointer_t<decltype(block)>;
            Workspace<T> ws(op, b.get(), 0.0, 0, stats, ctx->ctx, ve);
            if constexpr (sizeof(T) == 4) ws.set_fp32_rank_factor(fp32_rank_factor);
            const auto rr = ws.rank_reveal_block(block, ldx, w, max_rank);
            *rank = rr.rank;
            *cutoff = rr.cutoff;
            *discarded = rr.discarded;
            sv = rr.singular_values;
        };
        if (dtype == AMFLS_F64) run(static_cast<double*>(x));
        else run(static_cast<float*>(x));
        for (int i = 0; i < w; ++i) singular_values[i] = i < static_cast<int>(sv.size()) ? sv[i] : 0.0;
    });
}

int amfls_block_stats(amfls_context* ctx, int64_t rows, int32_t cols, const double* x, int64_t ldx, double* out) {
    return guard([&] {
        require(ctx && out && (rows == 0 || x), "null argument");
        require(rows >= 0 && cols >= 1 && ldx >= cols, "invalid block shape");
        ctx->ctx.block_stats<double>(rows, cols, x, ldx, out);
    });
}

}  // extern "C"
