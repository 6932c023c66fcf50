// Device operators of the AMFLS HIP engine.
//
// Counterpart of the reference's MatrixOperator contract
// (include/amfls/matrix_operator.hpp:29-89) for matrices that live in HBM:
// Y = A X and X = A^T Y on row-major device blocks, the relative block cost
// used by the width controller, and the normwise validation model
// {||A|| upper, ||A|| lower, apply steps, transpose steps} computed from the
// stored values exactly as the reference's constructors do
// (sparse_csr_operator.cpp:14-149, dense_operator.cpp:20-115,
// row_major_dense_operator.cpp:23-101).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <vector>

#include "context.hpp"

namespace amflsx {

struct ValidationModel {
    double upper = kInfD;
    double lower = 0.0;
    long long steps_apply = 0;
    long long steps_transpose = 0;
};

class DeviceOperator {
public:
    virtual ~DeviceOperator() = default;
    int64_t rows() const { return rows_; }
    int64_t cols() const { return cols_; }

    virtual void apply(hipStream_t s, const double* X, int ncols, int64_t ldx, double* Y, int64_t ldy) const = 0;
    virtual void apply_t(hipStream_t s, const double* Y, int ncols, int64_t ldy, double* X, int64_t ldx) const = 0;
    virtual void apply(hipStream_t s, const float* X, int ncols, int64_t ldx, float* Y, int64_t ldy) const = 0;
    virtual void apply_t(hipStream_t s, const float* Y, int ncols, int64_t ldy, float* X, int64_t ldx) const = 0;

    virtual double relative_block_product_cost(int s) const {
        return s > 0 ? static_cast<double>(s) : kInfD;
    }
    virtual ValidationModel validation_model() const { return {}; }

    // Long-double validation refinement on the host (reference
    // matrix_operator.hpp:62-88).  Supported where a host copy is affordable.
    virtual bool supports_refinement() const { return false; }
    virtual bool refine_apply(const long double*, long double*) const { return false; }
    virtual bool refine_apply_t(const long double*, long double*) const { return false; }

    // Algorithmic HBM bytes of one product of width s in the given working
    // precision (4 or 8 bytes per block entry).
    virtual double product_bytes(int s, bool transpose, int block_elem_bytes) const = 0;
    virtual double product_flops(int s) const = 0;

    // Selectable kernel variants (CSR SpMM); -1 = default.
    virtual void set_variant(int, bool, int) {}

    // The same product as apply / apply_t on FP64 (fp64 = true) or FP32
    // blocks, enqueued on `s` with `scratch` (a fixed, non-growing buffer)
    // as the only device memory beyond the blocks: no allocation, no timing
    // events or stamps.  Returns false for operators without such a path.
    virtual bool product_noalloc(hipStream_t, bool /*transpose*/, const void* /*in*/, int /*ncols*/,
                                 int64_t /*ld_in*/, void* /*out*/, int64_t /*ld_out*/, bool /*fp64*/,
                                 Scratch& /*scratch*/) const {
        return false;
    }

    // Operators whose product kernels time themselves: stamp_next() hands the
    // next apply / apply_t the device values its kernels stamp their start
    // and end into (ProductStamp); the product consumes them.  The hand-off
    // is per host thread (every caller sets it and runs the product on the
    // same thread), so workers that share one operator cannot swap stamps.
    virtual bool times_own_kernels() const { return false; }
    void stamp_next(unsigned long long* slots) const { stamp_ = slots; }
    void clear_stamp() const { stamp_ = nullptr; }

    // Kernel timing with HIP events recorded on the launching stream around
    // every product (off by default).  Stats: launches, seconds, algorithmic
    // bytes (product_bytes of each launch).  The event queues are locked; the
    // begin/end pairing assumes one thread drives the operator while it is on.
    void set_profiling(bool on);
    bool profiling() const { return profiling_.load(); }
    void kernel_stats(long long* launches, double* seconds, double* bytes);
    // The same totals over the products of one class (direction, block
    // widths in [min_cols, max_cols], block element bytes) since the last
    // kernel_stats; does not reset.
    void kernel_stats_class(bool transpose, int min_cols, int max_cols, int elem_bytes, long long* launches,
                            double* seconds, double* bytes);

protected:
    void prof_begin(hipStream_t s) const;
    void prof_end(hipStream_t s, int ncols, bool transpose, int elem_bytes) const;
    // Events for kernels that take their dispatch timestamps (nullptr when
    // profiling is off), and the record of one product timed that way.
    void prof_events(hipEvent_t* start, hipEvent_t* stop) const;
    void prof_commit(hipEvent_t start, hipEvent_t stop, int ncols, bool transpose, int elem_bytes) const;
    unsigned long long* take_stamp() const {
        unsigned long long* st = stamp_;
        stamp_ = nullptr;
        return st;
    }
    int64_t rows_ = 0, cols_ = 0;

private:
    static inline thread_local unsigned long long* stamp_ = nullptr;
    struct ProfPair { hipEvent_t start, stop; double bytes; int ncols, elem_bytes; bool transpose; };
    struct ProfRecord { double seconds, bytes; int ncols, elem_bytes; bool transpose; };
    mutable std::mutex prof_mu_;  // products of one operator may run on several threads
    std::atomic<bool> profiling_{false};
    mutable std::vector<ProfPair> pending_;
    std::vector<ProfRecord> records_;
    mutable std::vector<hipEvent_t> pool_;
    mutable hipEvent_t open_start_ = nullptr;
    long long launches_ = 0;
    double seconds_ = 0.0, bytes_ = 0.0;
    void resolve_pending();
};

// Ends a stamp hand-off with the product that was to consume it: the
// destructor clears the thread's pending stamp whether the product returned
// or threw, so no stamp outlives its product (its slots may belong to a
// timer that is gone by the next product on this thread).
struct StampScope {
    const DeviceOperator& op;
    ~StampScope() { op.clear_stamp(); }
};

// Non-owning CSR (int64 row pointers, int32 column indices, fp32/fp64
// values) on the device; owns its device-built transpose.
class CsrOperator final : public DeviceOperator {
public:
    CsrOperator(int64_t rows, int64_t cols, int64_t nnz, const int64_t* row_ptr, const int32_t* col_idx,
                const void* values, ValueType type, Context& ctx);

    void apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const override;
    void apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const override;
    ValidationModel validation_model() const override { return model_; }
    bool supports_refinement() const override { return true; }
    bool refine_apply(const long double* x, long double* y) const override;
    bool refine_apply_t(const long double* y, long double* x) const override;
    double product_bytes(int s, bool transpose, int eb) const override;
    double product_flops(int s) const override { return 2.0 * static_cast<double>(nnz_) * s; }
    void set_variant(int s, bool transpose, int variant) override;
    bool product_noalloc(hipStream_t s, bool transpose, const void* in, int ncols, int64_t ld_in, void* out,
                         int64_t ld_out, bool fp64, Scratch& scratch) const override;
    bool times_own_kernels() const override { return true; }
    // Times the SpMM variants whose column lanes fit each block width (the
    // next power of two and twice it) on this matrix, with FP64 or FP32
    // blocks, and keeps the fastest for A and A^T (median kernel time of
    // `reps` launches).  Returns the chosen variants, (A, A^T) per width.
    std::vector<int> autotune(Context& ctx, const std::vector<int>& widths, bool fp64_blocks, int reps);
    // Builds the slab-tiled copy of A (or of A^T) that the width-one slab
    // kernel reads, unless it exists.  Without `force` it is built only where
    // staging x once per panel moves at most 8 times the entries' bytes (x is
    // reused enough within a panel).  Returns whether the copy exists.
    bool ensure_slab(bool transpose, bool force);

    // Augmented [A; a I] products: Y rows [rows, rows+cols) = a X; the
    // transpose adds a * Y[rows:] to A^T Y[0:rows].
    template <typename T>
    void apply_t_add(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx, const T* add,
                     int64_t ld_add, double a) const;

    const CsrView& view() const { return a_; }
    const CsrView& transpose_view() const { return at_; }
    int64_t nnz() const { return nnz_; }

private:
    int variant_for(int s, bool transpose) const;
    void ensure_host_copy() const;

    CsrView a_, at_;
    int64_t nnz_ = 0;
    DevBuf<int64_t> t_rp_, stream_a_, stream_at_, stream_a_wide_, stream_at_wide_;
    struct SlabCopy {
        DevBuf<int64_t> panels, tiles;
        DevBuf<uint32_t> idx;
        DevBuf<char> val;
    };
    SlabCopy slab_a_, slab_at_;
    DevBuf<int32_t> t_ci_;
    DevBuf<char> t_val_;
    ValidationModel model_;
    std::map<std::pair<int, bool>, int> variants_;
    mutable std::once_flag host_once_;
    mutable std::vector<int64_t> h_rp_;
    mutable std::vector<int32_t> h_ci_;
    mutable std::vector<double> h_val_;
};

// Non-owning dense A (fp64 / fp32 / bf16), column-major (lda >= rows) or
// row-major (lda >= cols).
class DenseOperator final : public DeviceOperator {
public:
    DenseOperator(int64_t rows, int64_t cols, const void* values, ValueType type, bool row_major,
                  int64_t ld, Context& ctx);

    void apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const override;
    void apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const override;
    double relative_block_product_cost(int s) const override { return s > 0 ? 1.0 : kInfD; }
    ValidationModel validation_model() const override { return model_; }
    bool supports_refinement() const override;
    bool refine_apply(const long double* x, long double* y) const override;
    bool refine_apply_t(const long double* y, long double* x) const override;
    double product_bytes(int s, bool transpose, int eb) const override;
    double product_flops(int s) const override {
        return 2.0 * static_cast<double>(rows_) * static_cast<double>(cols_) * s;
    }
    Strided strided() const { return a_; }
    // Row-major storage and blocks of at most 64 columns run on the matrix
    // cores; `reduced` makes fp64-block products fp32-accurate (split-bf16
    // or fp32 matrix cores) for the inner solves of mixed precision.
    bool use_mfma(int ncols) const { return mfma_ && a_.cs == 1 && ncols <= 64; }
    void set_mfma(bool on) { mfma_ = on; }
    void set_reduced_products(bool on) { reduced_ = on; }
    // Row-major storage and single-column blocks run on the vector ALUs
    // (FP64 accumulation) by default; widths up to 8 can be sent there with
    // set_variant(s, transpose, v): v = 0 the vector-ALU kernel, 1 the
    // matrix cores, -1 the default.
    void set_variant(int s, bool transpose, int variant) override;
    bool use_narrow(int ncols, bool transpose) const;
    bool product_noalloc(hipStream_t s, bool transpose, const void* in, int ncols, int64_t ld_in, void* out,
                         int64_t ld_out, bool fp64, Scratch& scratch) const override;

private:
    bool mfma_ = true;
    bool reduced_ = false;
    std::map<std::pair<int, bool>, int> variants_;
    template <typename T>
    void do_apply(hipStream_t s, const T* X, int n, int64_t ldx, T* Y, int64_t ldy, Scratch& scratch,
                  bool profile) const;
    template <typename T>
    void do_apply_t(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx, Scratch& scratch,
                    bool profile) const;
    // Streams A to the host in chunks of whole rows (at most 2^22 entries,
    // widened to double, row-major) and hands each to use(first_row, rows,
    // values); stops early when use returns false.  No copy of A is kept.
    bool stream_rows(const std::function<bool(int64_t, int64_t, const double*)>& use) const;

    Strided a_;
    ValidationModel model_;
    Context& ctx_;
};

// A caller's operator that lives on the host (the reference's custom
// MatrixOperator subclasses, matrix_operator.hpp:29-89): each product copies
// the engine's device block to page-locked host memory, hands it to the
// callback as a column-major FP64 block (x: cols x s, y: rows x s, leading
// dimension = rows of the block, matrix_operator.hpp:36-44), and copies the
// result back.  The callbacks run synchronously on the calling thread; a
// nonzero return is an error.  FP32 blocks are widened to FP64 for the
// callback and rounded back.
using HostApplyFn = int (*)(void* user, const double* in, int block_cols, double* out);
// Optional hooks of a caller's operator (reference matrix_operator.hpp:50-88):
// the scheduler's relative block-product cost (absent: s) and the
// long-double validation refinement products y = A x, x = A^T y (absent or
// either missing: no refinement), each returning 0 on success.
using HostCostFn = double (*)(void* user, int block_cols);
using HostRefineFn = int (*)(void* user, const long double* in, long double* out);
struct HostHooks {
    HostCostFn cost = nullptr;
    HostRefineFn refine = nullptr, refine_t = nullptr;
};

class HostOperator final : public DeviceOperator {
public:
    HostOperator(int64_t rows, int64_t cols, HostApplyFn apply, HostApplyFn apply_transpose, void* user,
                 ValidationModel model, HostHooks hooks = {});
    double relative_block_product_cost(int s) const override {
        return hooks_.cost ? hooks_.cost(user_, s) : DeviceOperator::relative_block_product_cost(s);
    }
    bool supports_refinement() const override { return hooks_.refine != nullptr && hooks_.refine_t != nullptr; }
    bool refine_apply(const long double* x, long double* y) const override {
        return supports_refinement() && hooks_.refine(user_, x, y) == 0;
    }
    bool refine_apply_t(const long double* y, long double* x) const override {
        return supports_refinement() && hooks_.refine_t(user_, y, x) == 0;
    }
    void apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const override;
    void apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const override;
    ValidationModel validation_model() const override { return model_; }
    // host-memory bytes the callback reads and writes
    double product_bytes(int s, bool, int) const override {
        return 8.0 * s * static_cast<double>(rows_ + cols_);
    }
    double product_flops(int) const override { return 0.0; }

private:
    template <typename T>
    void call(hipStream_t s, HostApplyFn fn, const T* in, int64_t in_rows, int n, int64_t ldi, T* out,
              int64_t out_rows, int64_t ldo) const;

    HostApplyFn apply_, apply_t_;
    void* user_;
    ValidationModel model_;
    HostHooks hooks_;
    mutable PinnedBuf<char> hin_, hout_;
    mutable std::vector<double> cin_, cout_;
};

// [A; a I] for the ridge correction problems of the mixed-precision outer
// refinement (rows = m + n).
class AugmentedOperator final : public DeviceOperator {
public:
    AugmentedOperator(const DeviceOperator& base, double a, Context& ctx);
    void apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const override;
    void apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const override;
    void apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const override;
    double relative_block_product_cost(int s) const override { return base_.relative_block_product_cost(s); }
    ValidationModel validation_model() const override { return model_; }
    double product_bytes(int s, bool transpose, int eb) const override {
        return base_.product_bytes(s, transpose, eb) + 2.0 * eb * s * static_cast<double>(base_.cols());
    }
    // a CSR base times its own kernels (the augmented rows' copy excepted)
    bool times_own_kernels() const override { return base_.times_own_kernels(); }
    double product_flops(int s) const override { return base_.product_flops(s); }

private:
    template <typename T>
    void do_apply(hipStream_t s, const T* X, int n, int64_t ldx, T* Y, int64_t ldy) const;
    template <typename T>
    void do_apply_t(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx) const;

    const DeviceOperator& base_;
    double a_;
    ValidationModel model_;
};

// Bounds from per-row / per-column |a| sums and square sums.
ValidationModel norm_model_from_sums(const std::vector<double>& row_sum, const std::vector<double>& row_sq,
                                     const std::vector<long long>& row_ops,
                                     const std::vector<double>& col_sum, const std::vector<double>& col_sq,
                                     const std::vector<long long>& col_ops, double max_entry,
                                     long long steps_apply, long long steps_transpose);

}  // namespace amflsx
