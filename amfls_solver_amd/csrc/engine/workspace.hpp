// Device block Golub-Kahan workspace (one variable-width process).
//
// Follows the reference's BlockGolubKahanWorkspace
// (src/algorithms/mathematics/krylov/krylov_workspace.cpp:335-925): V is
// retained in full, only the active left frontier U (plus pending injected
// columns P*Omega) is stored, T and g = U^T b grow on the host, and every
// right/left residual is block-CGS2 (or DGKS for one column) reorthogonalized
// and rank-revealed.  The big blocks live in HBM in row-major layout; the
// rank-revealing thin SVD of the reference becomes CholeskyQR2 (shifted
// when needed) on the device plus a host SVD of the small R factor.
#pragma once

#include <memory>
#include <vector>

#include "operator.hpp"
#include "small_la.hpp"

namespace amflsx {

struct StepAudit {
    int requested_injected_columns = 0, retained_injected_columns = 0, left_input_width = 0;
    int retained_right_width = 0, retained_next_left_width = 0, basis_rank_before = 0, basis_rank_after = 0;
    double right_rank_cutoff = 0.0, right_discarded_residual_norm = 0.0,
           right_active_discarded_residual_norm = 0.0, left_rank_cutoff = 0.0,
           left_discarded_residual_norm = 0.0;
    bool right_rank_loss = false, left_rank_loss = false, right_positive_finite_below_cutoff = false,
         left_positive_finite_below_cutoff = false, right_nonfinite_residual = false,
         left_nonfinite_residual = false, right_basis_cap_truncation = false, left_basis_cap_truncation = false;
};

bool step_has_numerical_rank_truncation(const StepAudit& s);
bool step_has_nonfinite_breakdown(const StepAudit& s);

// Growable row-major device block (rows fixed, capacity in columns).
template <typename T>
class DeviceBlock {
public:
    void init(int64_t rows, int cap_cols);
    void ensure_cols(int cols, hipStream_t s);  // grows capacity, preserving contents
    T* col(int j) const { return buf_.get() + j; }
    T* ptr() const { return buf_.get(); }
    int64_t ld() const { return ld_; }
    int64_t rows() const { return rows_; }
    int capacity() const { return static_cast<int>(ld_); }

private:
    DevBuf<T> buf_;
    int64_t rows_ = 0, ld_ = 0;
};

class ValidationEngine;

template <typename T>
class Workspace {
public:
    Workspace(const DeviceOperator& op, const double* b_dev, double regularization, int maximum_basis_size,
              RunStatistics& stats, Context& ctx, ValidationEngine& validator);
    ~Workspace() { ctx_.timer.discard(); }
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;

    void initialize(bool with_aux, int aux_cols, uint64_t seed, uint64_t stream, uint64_t first_col);
    int append_left_seed_gaussian(int cols, uint64_t seed, uint64_t stream, uint64_t first_col,
                                  bool orthogonal_to_rhs);
    bool advance_level();
    // Advance toward `target_depth`: while the frontier is one column wide
    // (no pending injected columns), levels are enqueued in batches whose
    // rank decisions are taken on the device and read back once per batch;
    // otherwise, or for a level the device declined, one host-driven level
    // (advance_level).  Returns false, like advance_level, when the last
    // level attempted did not extend the basis; the audit is that level's.
    bool advance_toward(int target_depth);
    LeastSquaresResult evaluate(double tolerance, int epoch, int auxiliary_width);
    // The regularization of the projected solve and the validation.  The
    // Golub-Kahan process itself does not depend on it, so one basis can be
    // evaluated at several values.
    void set_regularization(double lambda) { lambda_ = lambda; }
    // Replaces the rank-cutoff factor of FP32 blocks (cutoff = factor *
    // scale); 0 restores the default rule.
    void set_fp32_rank_factor(double f) { fp32_rank_factor_ = f; }
    double regularization() const { return lambda_; }
    // Candidate x = V y from the current projected solve (device, double).
    void candidate(double* x_dev);

    int depth() const { return depth_; }
    int basis_rank() const { return k_; }
    int basis_limit() const { return basis_limit_; }
    int active_left_width() const { return active_width_; }
    bool can_advance() const;
    double right_hand_side_norm() const { return bnorm_; }
    const StepAudit& last_step_audit() const { return audit_; }
    const Staircase& projected_operator() const { return T_; }
    // The retained right basis V: n x basis_rank(), row-major with leading
    // dimension basis_ld().
    const T* basis() const { return V_.ptr(); }
    int64_t basis_ld() const { return V_.ld(); }

    // Rank reveal of one block W (m x w, row-major, leading dimension ldw)
    // with no basis to orthogonalize against: the factorization every new
    // search block goes through (the cutoff scale is W's Frobenius norm).
    // W is overwritten with its orthonormal factor.
    struct RankReveal {
        int rank = 0;
        double cutoff = 0.0, discarded = 0.0;
        std::vector<double> singular_values;  // of the retained factor, decreasing
    };
    RankReveal rank_reveal_block(T* W, int64_t ldw, int w, int max_rank);

private:
    struct Factor {
        std::vector<double> coefficients;  // basis_cols x w (row-major), when retained
        int rank = 0;
        HostMat factor;                     // rank x w
        double discarded = 0.0, discarded_tracked = 0.0, cutoff = 0.0;
        bool positive_finite_below_cutoff = false, nonfinite_residual = false;
    };
    // Orthogonalize W (N x w, in place) against basis (N x k) and rank-reveal.
    // On return W's first `rank` columns hold the orthonormal q.  The scale
    // of the rank cutoff is the norm whose statistics the caller enqueued
    // into slot 0 (kScaleSlot) before W was modified.
    Factor orthogonalize_and_factor(T* W, int64_t ldw, int64_t N, int w, const T* basis, int64_t ldb, int k,
                                    int max_rank, int tracked, bool retain);
    // Device statistics slots (triples) of one orthogonalization and the
    // DGKS gate; copied to the host with a single wait.
    static constexpr int kScaleSlot = 0, kBeforeSlot = 3, kFirstSlot = 6, kSecondSlot = 9, kCoeffSlot = 12,
                         kGateSlot = 15, kSlotCount = 16;
    void enqueue_scale(const T* X, int64_t ldx, int64_t rows, int cols);
    // Device-resident rank-one levels.
    int fast_batch_limit() const;
    int run_fast_batch(int count);  // returns levels committed
    void enqueue_product(const T* X, int64_t ldx, T* Y, int64_t ldy, bool transpose);
    // Wall clock and an operator whose kernels time themselves: arms the
    // operator's next product to add its kernel time to `seconds`.
    bool time_with_stamp(double* seconds);
    double rank_factor(int k, int w) const;  // rank cutoff = factor * scale
    void sync_projected_rhs();
    double block_norm(const T* X, int64_t ldx, int64_t rows, int cols);
    double rank_cutoff(double scale, int k, int w) const;
    void apply_search(const T* X, int s, int64_t ldx, T* Y, int64_t ldy, bool transpose);
    void ensure_basis(int cols);
    void small_multiply(T* W, int64_t ldw, int64_t N, int w, int r, const HostMat& M);

    const DeviceOperator& op_;
    Context& ctx_;
    RunStatistics& stats_;
    ValidationEngine& validator_;
    int64_t m_, n_;
    const double* b_dev_;
    double bnorm_ = 0.0;
    double lambda_;
    int basis_limit_;
    DeviceBlock<T> V_;       // n x k
    DeviceBlock<T> left_;    // m x (a + p): active U then pending P*Omega
    DeviceBlock<T> work_;    // m x r: next left residual / q
    DeviceBlock<T> right_;   // n x (a + p): right residual
    DevBuf<double> coeff_;   // k x w FP64 coefficient accumulation
    DevBuf<double> cbuf_;    // pass coefficients
    DevBuf<double> gdev_;    // g = U^T b entries computed on the device
    DevBuf<double> xcand_;   // candidate solution of the current checkpoint (n)
    DevBuf<double> slots_;   // orthogonalization statistics (kSlotCount doubles)
    DevBuf<double> fast_cells_;  // go, gate, inverse norm, then one coupling cell per batched level
    DevBuf<double> fast_rec_;    // 8 doubles per batched level (right and left decisions)
    DevBuf<double> fast_part_;   // per-block partials of the level chain
    std::vector<double> fast_host_, fast_gpart_;
    DevBuf<double> gram_;    // w x w Gram matrices
    DevBuf<T> tmp_;          // wide row-local products
    int k_ = 0;
    int active_first_ = 0, active_width_ = 0, pending_ = 0;
    int prev_first_ = 0, prev_width_ = 0;
    HostMat coupling_;       // active_width x prev_width
    Staircase T_;
    std::vector<double> g_;
    size_t g_host_ = 0;      // leading entries of g_ already copied from gdev_
    int pending_requested_ = 0, pending_retained_ = 0;
    int depth_ = 0;
    double fp32_rank_factor_ = 0.0;
    bool initialized_ = false, reliable_ = true;
    StepAudit audit_;
    std::vector<double> y_;  // last projected solution
};

}  // namespace amflsx
