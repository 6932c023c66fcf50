#include "workspace.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "validation.hpp"

namespace amflsx {

bool step_has_numerical_rank_truncation(const StepAudit& s) {
    return (s.right_rank_loss && s.right_positive_finite_below_cutoff && s.right_active_discarded_residual_norm > 0.0) ||
           (s.left_rank_loss && s.left_positive_finite_below_cutoff);
}

bool step_has_nonfinite_breakdown(const StepAudit& s) {
    return s.right_nonfinite_residual || s.left_nonfinite_residual;
}

template <typename T>
void DeviceBlock<T>::init(int64_t rows, int cap_cols) {
    rows_ = rows;
    ld_ = std::max(1, cap_cols);
    buf_.reset(static_cast<size_t>(rows_ * ld_));
}

template <typename T>
void DeviceBlock<T>::ensure_cols(int cols, hipStream_t s) {
    if (cols <= ld_) return;
    int64_t nld = std::max<int64_t>(cols, 2 * ld_);
    DevBuf<T> nb(static_cast<size_t>(rows_ * nld));
    copy_block<T, T>(s, rows_, static_cast<int>(ld_), buf_.get(), ld_, nb.get(), nld, 1.0);
    AMFLS_HIP_CHECK(hipStreamSynchronize(s));
    buf_ = std::move(nb);
    ld_ = nld;
}

namespace {
constexpr ValueType vt_of(float) { return ValueType::f32; }
constexpr ValueType vt_of(double) { return ValueType::f64; }
template <typename T> constexpr double unit_roundoff() { return 0.5 * std::numeric_limits<T>::epsilon(); }
}  // namespace

template <typename T>
Workspace<T>::Workspace(const DeviceOperator& op, const double* b_dev, double regularization,
                        int maximum_basis_size, RunStatistics& stats, Context& ctx, ValidationEngine& validator)
    : op_(op), ctx_(ctx), stats_(stats), validator_(validator), m_(op.rows()), n_(op.cols()), b_dev_(b_dev),
      lambda_(regularization) {
    if (m_ <= 0 || n_ <= 0) throw std::invalid_argument("block Golub-Kahan operator dimensions must be positive");
    if (!std::isfinite(regularization) || regularization < 0.0)
        throw std::invalid_argument("block Golub-Kahan regularization must be finite and nonnegative");
    if (maximum_basis_size < 0) throw std::invalid_argument("block Golub-Kahan basis limit must be nonnegative");
    const int64_t mn = std::min(m_, n_);
    const int64_t lim = maximum_basis_size > 0 ? std::min<int64_t>(maximum_basis_size, mn) : mn;
    if (lim >= std::numeric_limits<int>::max()) throw std::length_error("basis limit exceeds int range");
    basis_limit_ = static_cast<int>(lim);
    bnorm_ = validator_.rhs_norm();
    slots_.reset(kSlotCount);
    gdev_.reset(static_cast<size_t>(basis_limit_) + 2);
    V_.init(n_, std::min(basis_limit_, 64));
    left_.init(m_, 1);
    work_.init(m_, 1);
    right_.init(n_, 1);
    T_.rows = 0;
}

template <typename T>
bool Workspace<T>::can_advance() const {
    return initialized_ && reliable_ && active_width_ + pending_ > 0 && k_ < basis_limit_;
}

template <typename T>
double Workspace<T>::block_norm(const T* X, int64_t ldx, int64_t rows, int cols) {
    double st[3];
    ctx_.block_stats<T>(rows, cols, X, ldx, st);
    return fp::nominal_norm(st[0], st[1], st[2]);
}

template <typename T>
double Workspace<T>::rank_factor(int k, int w) const {
    // reference krylov_workspace.cpp:63-80 for binary64 blocks.  FP32 blocks
    // carry a rounding floor near u_32 * ||block||, so the cutoff is scaled
    // by the working precision instead of the problem dimension.
    if constexpr (sizeof(T) == 8) {
        const double dim = static_cast<double>(std::max<int64_t>({1, m_, n_, static_cast<int64_t>(k),
                                                                  static_cast<int64_t>(w)}));
        return 64.0 * std::numeric_limits<double>::epsilon() * dim;
    } else {
        if (fp32_rank_factor_ > 0.0) return fp32_rank_factor_;
        const double dim = static_cast<double>(std::max({1, k, w}));
        return 64.0 * std::numeric_limits<float>::epsilon() * std::sqrt(dim);
    }
}

template <typename T>
double Workspace<T>::rank_cutoff(double scale, int k, int w) const {
    if (!(scale >= 0.0) || !std::isfinite(scale)) return kInfD;
    return rank_factor(k, w) * scale;
}

template <typename T>
void Workspace<T>::apply_search(const T* X, int s, int64_t ldx, T* Y, int64_t ldy, bool transpose) {
    // Wall clock: the product is bracketed by events on the stream and its
    // GPU time is added to the operator seconds when the timer is flushed.
    // Synthetic clock: the charged cost is the product's time.
    Clock& clk = ctx_.clock;
    const bool wall = !clk.synthetic();
    double* seconds = transpose ? &stats_.at_seconds : &stats_.a_seconds;
    const double charged_from = clk.now();
    clk.charge_operator(op_.relative_block_product_cost(s));
    if (!wall) *seconds += clk.now() - charged_from;
    const bool events = wall && !time_with_stamp(seconds);
    const StampScope stamp_scope{op_};
    if (events) ctx_.timer.begin(ctx_.stream);
    if (transpose) op_.apply_t(ctx_.stream, X, s, ldx, Y, ldy);
    else op_.apply(ctx_.stream, X, s, ldx, Y, ldy);
    if (events) ctx_.timer.end(ctx_.stream, seconds);
    if (transpose) {
        stats_.at_columns += s;
        stats_.search_at_columns += s;
        ++stats_.at_block_calls;
        ++stats_.search_at_block_calls;
    } else {
        stats_.a_columns += s;
        stats_.search_a_columns += s;
        ++stats_.a_block_calls;
        ++stats_.search_a_block_calls;
    }
}

template <typename T>
void Workspace<T>::enqueue_scale(const T* X, int64_t ldx, int64_t rows, int cols) {
    ctx_.block_stats_async<T>(rows, cols, X, ldx, slots_.get() + kScaleSlot);
}

template <typename T>
void Workspace<T>::sync_projected_rhs() {
    if (g_host_ >= g_.size()) return;
    ctx_.download(g_.data() + g_host_, gdev_.get() + g_host_, g_.size() - g_host_);
    g_host_ = g_.size();
}

template <typename T>
typename Workspace<T>::Factor Workspace<T>::orthogonalize_and_factor(T* W, int64_t ldw, int64_t N, int w,
                                                                     const T* basis, int64_t ldb, int k,
                                                                     int max_rank, int tracked, bool retain) {
    // reference krylov_workspace.cpp:82-310.  The projections, the DGKS
    // decision for one column and every norm run on the device; the host
    // waits once for the statistics (and retained coefficients) before the
    // rank decision.
    Clock& clk = ctx_.clock;
    const double start = clk.now();
    clk.charge_local(w, k);
    hipStream_t s = ctx_.stream;
    const Strided B{basis, vt_of(T{}), ldb, 1};
    double* sl = slots_.get();
    Factor f;
    if (k > 0) {
        coeff_.ensure(static_cast<size_t>(k) * w);
        cbuf_.ensure(static_cast<size_t>(k) * w);
        fill<double>(s, coeff_.get(), static_cast<int64_t>(k) * w, 0.0);
    }
    auto pass = [&](const double* gate) {
        // c = basis^T W ; W -= basis c ; coeff += c
        gemm_tn<T>(s, N, k, w, B, W, ldw, cbuf_.get(), ValueType::f64, w, 1.0, ctx_.scratch, gate);
        gemm_nn<T, double>(s, N, w, k, B, cbuf_.get(), w, W, ldw, -1.0, 1.0, gate);
        add_vec(s, static_cast<int64_t>(k) * w, cbuf_.get(), coeff_.get(), gate);
    };
    const bool dgks = w == 1 && k > 0;
    if (dgks) {
        // one-column DGKS with the fused rank-one kernels: second pass when
        // ||W|| fell below ||W_before|| / sqrt(2)
        unsigned* ctr = ctx_.counter.get();
        ctx_.block_stats_async<T>(N, 1, W, ldw, sl + kBeforeSlot);
        rank_one_project<T, T>(s, N, k, basis, ldb, W, ldw, cbuf_.get(), coeff_.get(), false, ctx_.scratch, ctr);
        rank_one_update<T>(s, N, k, basis, ldb, cbuf_.get(), W, ldw, sl + kFirstSlot, ctx_.scratch, ctr, nullptr,
                           nullptr, sl + kBeforeSlot, std::sqrt(0.5), sl + kGateSlot);
        rank_one_project<T, T>(s, N, k, basis, ldb, W, ldw, cbuf_.get(), coeff_.get(), true, ctx_.scratch, ctr,
                               sl + kGateSlot);
        rank_one_update<T>(s, N, k, basis, ldb, cbuf_.get(), W, ldw, sl + kSecondSlot, ctx_.scratch, ctr,
                           sl + kGateSlot);
    } else {
        for (int p = 0; p < (k > 0 ? 2 : 0); ++p) pass(nullptr);
        ctx_.block_stats_async<T>(N, w, W, ldw, sl + kFirstSlot);
    }
    if (k > 0) ctx_.block_stats_async<double>(static_cast<int64_t>(k) * w, 1, coeff_.get(), 1, sl + kCoeffSlot);
    const size_t ncoef = (retain && k > 0) ? static_cast<size_t>(k) * w : 0;
    double* land = reinterpret_cast<double*>(ctx_.landing.get());
    const bool coef_landing = (kSlotCount + ncoef) * sizeof(double) <= Context::kLandingBytes;
    AMFLS_HIP_CHECK(hipMemcpyAsync(land, sl, kSlotCount * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ncoef > 0 && coef_landing)
        AMFLS_HIP_CHECK(hipMemcpyAsync(land + kSlotCount, coeff_.get(), ncoef * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
    ctx_.sync();
    auto norm_at = [&](int slot) { return fp::nominal_norm(land[slot], land[slot + 1], land[slot + 2]); };
    const double scale = norm_at(kScaleSlot);
    double resid = norm_at(kFirstSlot);
    if (dgks && land[kGateSlot] != 0.0) resid = norm_at(kSecondSlot);
    const bool coeff_finite = k == 0 || land[kCoeffSlot + 2] == 0.0;
    if (ncoef > 0) {
        f.coefficients.resize(ncoef);
        if (coef_landing) std::copy(land + kSlotCount, land + kSlotCount + ncoef, f.coefficients.begin());
        else ctx_.download(f.coefficients.data(), coeff_.get(), ncoef);
    }
    f.cutoff = rank_cutoff(scale, k, w);
    f.discarded = resid;
    const bool finite = std::isfinite(scale) && std::isfinite(resid) && coeff_finite;
    f.factor = HostMat(0, w);
    auto done = [&] { stats_.orthogonalization_seconds += clk.now() - start; };
    if (!finite) {
        f.nonfinite_residual = true;
        f.discarded = kInfD;
        f.discarded_tracked = tracked > 0 ? kInfD : 0.0;
        done();
        return f;
    }
    if (max_rank == 0 || w == 0) {
        double t = 0.0;
        for (int c = 0; c < tracked; ++c) t = std::hypot(t, block_norm(W + c, ldw, N, 1));
        f.discarded_tracked = t;
        done();
        return f;
    }
    if (w == 1) {
        if (std::isfinite(resid) && resid > f.cutoff) {
            copy_block<T, T>(s, N, 1, W, ldw, W, ldw, 1.0 / resid);
            f.rank = 1;
            f.factor = HostMat(1, 1);
            f.factor(0, 0) = resid;
            f.discarded = 0.0;
            f.discarded_tracked = 0.0;
        } else if (std::isfinite(resid) && resid > 0.0 && resid <= f.cutoff && max_rank > 0) {
            f.positive_finite_below_cutoff = true;
            if (tracked == 1) f.discarded_tracked = resid;
        }
        done();
        return f;
    }
    // Rank-revealing factorization W = Q R: CholeskyQR passes until Q^T Q is
    // the identity to working precision (a diagonal shift keeps the first
    // Cholesky defined for numerically rank-deficient blocks), then a host
    // SVD of the small R gives the singular values of W.
    // u: the storage precision of W, which sets the orthogonality target.
    // The Gram matrix is accumulated in FP64 from exact products of the
    // stored entries and the Cholesky runs in FP64 for both block types, so
    // when the Cholesky needs a shift, and how large it must be, is an FP64
    // question (ug); an FP32-sized shift would keep FP32 blocks of moderate
    // condition from converging in the allotted passes.
    const double u = unit_roundoff<T>(), ug = unit_roundoff<double>();
    HostMat Rtot(w, w);
    for (int i = 0; i < w; ++i) Rtot(i, i) = 1.0;
    std::vector<double> G(static_cast<size_t>(w) * w);
    gram_.ensure(static_cast<size_t>(w) * w);
    double* gd = gram_.get();
    const double orth_tol = 64.0 * w * u;
    for (int p = 0; p < 4; ++p) {
        gemm_tn<T>(s, N, w, w, Strided{W, vt_of(T{}), ldw, 1}, W, ldw, gd, ValueType::f64, w, 1.0, ctx_.scratch);
        ctx_.download(G.data(), gd, G.size());
        HostMat Gm(w, w);
        Gm.a = G;
        double trace = 0.0, dev = 0.0;
        for (int i = 0; i < w; ++i) trace += Gm(i, i);
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < w; ++j) dev = std::max(dev, std::abs(Gm(i, j) - (i == j ? 1.0 : 0.0)));
        if (!(trace > 0.0) || !std::isfinite(trace)) {
            if (p == 0) Rtot = HostMat(w, w);  // exactly zero block: rank 0 below
            break;
        }
        if (p > 0 && dev <= orth_tol) break;
        HostMat R;
        double ratio = 0.0;
        bool ok = cholesky_upper(Gm, R, &ratio);
        if (!ok || ratio < 64.0 * std::sqrt(ug * static_cast<double>(w))) {
            const double shift = 11.0 * (static_cast<double>(w) * (w + 1) * ug +
                                         (static_cast<double>(N) * w + static_cast<double>(w) * (w + 1)) * ug) *
                                 trace;
            HostMat Gs = Gm;
            for (int i = 0; i < w; ++i) Gs(i, i) += std::max(shift, std::numeric_limits<double>::min());
            ok = cholesky_upper(Gs, R, &ratio);
            if (!ok) break;
        }
        HostMat Rinv = upper_inverse(R);
        small_multiply(W, ldw, N, w, w, Rinv);
        Rtot = matmul(R, Rtot);
    }
    // A block whose triangular factor provably keeps every column needs no
    // SVD: sigma_min(R) >= 1 / ||R^{-1}||_F, so when that bound clears the
    // cutoff the numerical rank is w, nothing is discarded, and W = Q R is
    // already a rank-revealing factorization (Q spans the same columns as
    // the singular vectors; the block recurrence and the projected problem
    // are invariant under that rotation).  Near-deficient blocks, and blocks
    // capped below their width, take the SVD.
    if (w <= max_rank) {
        const HostMat Rinv = upper_inverse(Rtot);
        double fro = 0.0;
        for (double v : Rinv.a) fro += v * v;
        if (std::isfinite(fro) && fro > 0.0 && 1.0 / std::sqrt(fro) > f.cutoff) {
            f.factor = Rtot;
            f.rank = w;
            f.discarded = 0.0;
            f.discarded_tracked = 0.0;
            done();
            return f;
        }
    }
    Svd sv = jacobi_svd(Rtot);
    int rank = 0;
    while (rank < static_cast<int>(sv.s.size()) && rank < max_rank && std::isfinite(sv.s[rank]) &&
           sv.s[rank] > f.cutoff)
        ++rank;
    if (rank < static_cast<int>(sv.s.size()) && rank < max_rank) {
        const double fd = sv.s[rank];
        f.positive_finite_below_cutoff = std::isfinite(fd) && fd > 0.0 && fd <= f.cutoff;
    }
    if (rank > 0) {
        HostMat Ur(w, rank);
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < rank; ++j) Ur(i, j) = sv.U(i, j);
        small_multiply(W, ldw, N, w, rank, Ur);
        f.factor = HostMat(rank, w);
        for (int r = 0; r < rank; ++r)
            for (int c = 0; c < w; ++c) f.factor(r, c) = sv.s[r] * sv.Vt(r, c);
    }
    f.rank = rank;
    f.discarded = 0.0;
    for (int i = rank; i < static_cast<int>(sv.s.size()); ++i) {
        f.discarded = std::hypot(f.discarded, sv.s[i]);
        double tw = 0.0;
        for (int c = 0; c < tracked; ++c) tw = std::hypot(tw, sv.Vt(i, c));
        f.discarded_tracked = std::hypot(f.discarded_tracked, sv.s[i] * tw);
    }
    done();
    return f;
}

template <typename T>
typename Workspace<T>::RankReveal Workspace<T>::rank_reveal_block(T* W, int64_t ldw, int w, int max_rank) {
    if (w <= 0 || ldw < w) throw std::invalid_argument("rank reveal needs a block of at least one column");
    ctx_.block_stats_async<T>(m_, w, W, ldw, slots_.get() + kScaleSlot);
    Factor f = orthogonalize_and_factor(W, ldw, m_, w, nullptr, 0, 0, max_rank, 0, false);
    RankReveal r;
    r.rank = f.rank;
    r.cutoff = f.cutoff;
    r.discarded = f.discarded;
    if (f.rank > 0) r.singular_values = jacobi_svd(f.factor).s;
    return r;
}

template <typename T>
void Workspace<T>::small_multiply(T* W, int64_t ldw, int64_t N, int w, int r, const HostMat& M) {
    // W[:, 0:r] = W[:, 0:w] * M, row by row (in place)
    double* md = ctx_.upload(M.a.data(), M.a.size());
    if (w <= 128 && r <= 128) {
        rowmul<T>(ctx_.stream, N, w, r, W, ldw, md, W, ldw);
        return;
    }
    tmp_.ensure(static_cast<size_t>(N) * r);
    gemm_nn<T, double>(ctx_.stream, N, r, w, Strided{W, vt_of(T{}), ldw, 1}, md, r, tmp_.get(), r, 1.0, 0.0);
    copy_block<T, T>(ctx_.stream, N, r, tmp_.get(), r, W, ldw, 1.0);
}

template <typename T>
void Workspace<T>::ensure_basis(int cols) {
    V_.ensure_cols(cols, ctx_.stream);
}

template <typename T>
void Workspace<T>::initialize(bool with_aux, int aux_cols, uint64_t seed, uint64_t stream, uint64_t first_col) {
    if (initialized_) throw std::logic_error("block Golub-Kahan workspace is already initialized");
    initialized_ = true;
    audit_ = {};
    Clock& clk = ctx_.clock;
    const double start = clk.now();
    if (std::isfinite(bnorm_) && bnorm_ > 0.0 && basis_limit_ > 0) {
        copy_block<double, T>(ctx_.stream, m_, 1, b_dev_, 1, left_.ptr(), left_.ld(), 1.0 / bnorm_);
        g_.push_back(bnorm_);
        g_host_ = g_.size();
        T_.rows = 1;
        active_first_ = 0;
        active_width_ = 1;
    }
    stats_.orthogonalization_seconds += clk.now() - start;
    if (bnorm_ > 0.0 && with_aux && aux_cols > 0)
        append_left_seed_gaussian(aux_cols, seed, stream, first_col, false);
}

template <typename T>
int Workspace<T>::append_left_seed_gaussian(int cols, uint64_t seed, uint64_t stream, uint64_t first_col,
                                            bool orthogonal_to_rhs) {
    // reference krylov_workspace.cpp:505-587 plus the generation of the
    // Gaussian block (amfls.cpp:191-231, gaussian.cpp:39-55)
    if (!initialized_) throw std::logic_error("block Golub-Kahan workspace is not initialized");
    if (!reliable_) throw std::logic_error("cannot continue block Golub-Kahan after numerical truncation");
    if (pending_requested_ == 0 && pending_ == 0) audit_ = {};
    pending_requested_ += cols;
    audit_.requested_injected_columns = pending_requested_;
    audit_.retained_injected_columns = pending_retained_;
    if (cols == 0) return 0;
    hipStream_t s = ctx_.stream;
    const int a = active_width_, p = pending_;
    left_.ensure_cols(a + p + cols, s);
    T* seed_ptr = left_.col(a + p);
    // Generate in FP64 so the projection off b-hat matches the reference
    // arithmetic, then store in the working precision.
    DevBuf<double> rnd(static_cast<size_t>(m_) * cols);
    gaussian_block<double>(s, m_, cols, seed, stream, first_col, rnd.get(), cols);
    if (orthogonal_to_rhs && bnorm_ != 0.0) {
        DevBuf<double> bhat(static_cast<size_t>(m_));
        copy_block<double, double>(s, m_, 1, b_dev_, 1, bhat.get(), 1, 1.0 / bnorm_);
        DevBuf<double> coef(static_cast<size_t>(cols));
        gemm_tn<double>(s, m_, cols, 1, Strided{rnd.get(), ValueType::f64, cols, 1}, bhat.get(), 1, coef.get(),
                        ValueType::f64, 1, 1.0, ctx_.scratch);
        // rnd(:, c) -= coef[c] * bhat  ==  rnd -= bhat (m x 1) * coef^T (1 x cols)
        gemm_nn<double, double>(s, m_, cols, 1, Strided{bhat.get(), ValueType::f64, 1, 1}, coef.get(), cols,
                                rnd.get(), cols, -1.0, 1.0);
    }
    copy_block<double, T>(s, m_, cols, rnd.get(), cols, seed_ptr, left_.ld(), 1.0);
    const int avail = std::max(0, basis_limit_ - k_ - a - p);
    const int max_rank = static_cast<int>(std::min<int64_t>({static_cast<int64_t>(avail), m_ - a - p,
                                                             static_cast<int64_t>(cols)}));
    if (max_rank <= 0) return 0;
    enqueue_scale(seed_ptr, left_.ld(), m_, cols);
    Factor f = orthogonalize_and_factor(seed_ptr, left_.ld(), m_, cols, left_.ptr(), left_.ld(), a + p, max_rank,
                                        0, false);
    if (f.rank <= 0) return 0;
    pending_ += f.rank;
    pending_retained_ += f.rank;
    audit_.retained_injected_columns = pending_retained_;
    return f.rank;
}

template <typename T>
bool Workspace<T>::advance_level() {
    // reference krylov_workspace.cpp:589-831
    StepAudit a;
    a.requested_injected_columns = pending_requested_;
    a.retained_injected_columns = pending_retained_;
    a.basis_rank_before = a.basis_rank_after = k_;
    a.left_input_width = active_width_ + pending_;
    audit_ = a;
    if (!can_advance()) return false;
    if (active_first_ + active_width_ != static_cast<int>(g_.size()))
        throw std::logic_error("block Golub-Kahan active left frontier is inconsistent");
    hipStream_t s = ctx_.stream;
    const int old_right = k_;
    const int old_left = static_cast<int>(g_.size());
    const int aw = active_width_;
    const int cw = aw + pending_;
    right_.ensure_cols(cw, s);
    apply_search(left_.ptr(), cw, left_.ld(), right_.ptr(), right_.ld(), true);
    pending_ = 0;
    pending_requested_ = pending_retained_ = 0;
    enqueue_scale(right_.ptr(), right_.ld(), n_, cw);
    if (prev_width_ > 0 && aw > 0) {
        if (coupling_.rows != aw || coupling_.cols != prev_width_)
            throw std::logic_error("block Golub-Kahan frontier coupling dimensions are invalid");
        // R -= V[:, prev] * cc^T with cc = [coupling; 0] (cw x prev_width)
        HostMat cct(prev_width_, cw);
        for (int i = 0; i < aw; ++i)
            for (int j = 0; j < prev_width_; ++j) cct(j, i) = coupling_(i, j);
        double* cd = ctx_.upload(cct.a.data(), cct.a.size());
        gemm_nn<T, double>(s, n_, cw, prev_width_, Strided{V_.col(prev_first_), vt_of(T{}), V_.ld(), 1}, cd, cw,
                           right_.ptr(), right_.ld(), -1.0, 1.0);
    }
    const int max_right = std::min(cw, basis_limit_ - k_);
    Factor rf = orthogonalize_and_factor(right_.ptr(), right_.ld(), n_, cw, V_.ptr(), V_.ld(), k_, max_right, aw,
                                         false);
    const int rw = rf.rank;
    audit_.retained_right_width = rw;
    audit_.right_rank_cutoff = rf.cutoff;
    audit_.right_discarded_residual_norm = rf.discarded;
    audit_.right_active_discarded_residual_norm = rf.discarded_tracked;
    audit_.right_positive_finite_below_cutoff = rf.positive_finite_below_cutoff;
    audit_.right_nonfinite_residual = rf.nonfinite_residual;
    audit_.right_rank_loss = rw < max_right;
    audit_.right_basis_cap_truncation = max_right < cw;
    if (step_has_numerical_rank_truncation(audit_) || step_has_nonfinite_breakdown(audit_)) reliable_ = false;
    if (rw == 0) {
        active_first_ = old_left;
        active_width_ = 0;
        prev_first_ = k_;
        prev_width_ = 0;
        coupling_ = HostMat();
        return false;
    }
    // L = A * newV ; L -= U_active * coup^T
    work_.ensure_cols(rw, s);
    apply_search(right_.ptr(), rw, right_.ld(), work_.ptr(), work_.ld(), false);
    enqueue_scale(work_.ptr(), work_.ld(), m_, rw);
    if (aw > 0) {
        HostMat ct(aw, rw);  // coup^T, coup = rf.factor[:, :aw] (rw x aw)
        for (int i = 0; i < rw; ++i)
            for (int j = 0; j < aw; ++j) ct(j, i) = rf.factor(i, j);
        double* cd = ctx_.upload(ct.a.data(), ct.a.size());
        gemm_nn<T, double>(s, m_, rw, aw, Strided{left_.ptr(), vt_of(T{}), left_.ld(), 1}, cd, rw, work_.ptr(),
                           work_.ld(), -1.0, 1.0);
    }
    const int max_left = static_cast<int>(std::min<int64_t>({static_cast<int64_t>(rw), m_ - old_left,
                                                             static_cast<int64_t>(basis_limit_) + 1 - old_left}));
    Factor lf = orthogonalize_and_factor(work_.ptr(), work_.ld(), m_, rw, left_.ptr(), left_.ld(), aw, max_left, 0,
                                         true);
    const int lw = lf.rank;
    audit_.retained_next_left_width = lw;
    audit_.left_rank_cutoff = lf.cutoff;
    audit_.left_discarded_residual_norm = lf.discarded;
    audit_.left_positive_finite_below_cutoff = lf.positive_finite_below_cutoff;
    audit_.left_nonfinite_residual = lf.nonfinite_residual;
    audit_.left_rank_loss = lw < max_left;
    audit_.left_basis_cap_truncation = max_left < rw;
    if (step_has_numerical_rank_truncation(audit_) || step_has_nonfinite_breakdown(audit_)) reliable_ = false;
    if (lf.nonfinite_residual) return false;
    // append V
    ensure_basis(k_ + rw);
    copy_block<T, T>(s, n_, rw, right_.ptr(), right_.ld(), V_.col(k_), V_.ld(), 1.0);
    k_ += rw;
    // T columns
    const int rows_after = old_left + lw;
    for (int c = 0; c < rw; ++c) {
        std::vector<double> col(static_cast<size_t>(rows_after - active_first_), 0.0);
        for (int r = 0; r < aw; ++r)
            col[r] = rf.factor(c, r) + lf.coefficients[static_cast<size_t>(r) * rw + c];
        for (int r = 0; r < lw; ++r) col[old_left - active_first_ + r] = lf.factor(r, c);
        T_.lo.push_back(active_first_);
        T_.val.push_back(std::move(col));
    }
    T_.rows = rows_after;
    // new U = q (first lw columns of work_) -> left_
    std::swap(left_, work_);
    // g entries U_new^T b
    if (lw > 0) {
        // g entries stay on the device until the next projected solve
        if (g_.size() + lw > gdev_.size()) throw std::logic_error("projected right-hand side exceeds its capacity");
        cbuf_.ensure(static_cast<size_t>(lw));
        rank_one_project<T, double>(s, m_, lw, left_.ptr(), left_.ld(), b_dev_, 1, gdev_.get() + g_.size(),
                                    cbuf_.get(), false, ctx_.scratch, ctx_.counter.get());
        g_.resize(g_.size() + lw, std::numeric_limits<double>::quiet_NaN());
    }
    active_first_ = old_left;
    active_width_ = lw;
    prev_first_ = old_right;
    prev_width_ = rw;
    coupling_ = lf.factor;
    ++depth_;
    audit_.basis_rank_after = k_;
    return true;
}

template <typename T>
void Workspace<T>::enqueue_product(const T* X, int64_t ldx, T* Y, int64_t ldy, bool transpose) {
    // Counted and (on the synthetic clock) charged when the level commits.
    double* seconds = transpose ? &stats_.at_seconds : &stats_.a_seconds;
    const bool events = !ctx_.clock.synthetic() && !time_with_stamp(seconds);
    const StampScope stamp_scope{op_};
    if (events) ctx_.timer.begin(ctx_.stream);
    if (transpose) op_.apply_t(ctx_.stream, X, 1, ldx, Y, ldy);
    else op_.apply(ctx_.stream, X, 1, ldx, Y, ldy);
    if (events) ctx_.timer.end(ctx_.stream, seconds);
}

template <typename T>
bool Workspace<T>::time_with_stamp(double* seconds) {
    // wall clock with an operator that times its own kernels: the product
    // adds its kernel time to `seconds` through the context's device stamp
    if (ctx_.clock.synthetic() || !op_.times_own_kernels()) return false;
    op_.stamp_next(ctx_.timer.kernel_stamp(ctx_.stream, seconds));
    return true;
}

template <typename T>
int Workspace<T>::fast_batch_limit() const {
    // Frontier of width one with a 1 x 1 coupling and nothing pending: every
    // level of the batch then has max_right = max_left = 1.
    if (!initialized_ || !reliable_ || pending_ != 0 || active_width_ != 1 || prev_width_ != 1 || k_ < 1 ||
        coupling_.rows != 1 || coupling_.cols != 1 || active_first_ + 1 != static_cast<int>(g_.size()))
        return 0;
    const int64_t old_left = static_cast<int64_t>(g_.size());
    // the chain's update kernel keeps the k coefficients in LDS: k <= 2048
    const int64_t lim = std::min<int64_t>({static_cast<int64_t>(basis_limit_) - k_, m_ - old_left,
                                           static_cast<int64_t>(basis_limit_) + 1 - old_left, 64,
                                           static_cast<int64_t>(2048) - k_});
    return static_cast<int>(std::max<int64_t>(0, lim));
}

namespace {

// Stream capture of one batch of levels into a graph that is launched once
// (Context::level_graphs).  The graph is destroyed with this object, after
// the batch's results have been downloaded (the download waits for them).
class BatchGraph {
public:
    explicit BatchGraph(hipStream_t s) : s_(s) {
        if (s_ == nullptr) return;
        AMFLS_HIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
    }
    void launch() {
        if (s_ == nullptr) return;
        capturing_ = false;
        AMFLS_HIP_CHECK(hipStreamEndCapture(s_, &graph_));
        AMFLS_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        AMFLS_HIP_CHECK(hipGraphLaunch(exec_, s_));
    }
    ~BatchGraph() {
        if (capturing_) {  // an exception left the capture open: close it, drop what was captured
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        // a launched graph has finished before it is destroyed (the batch's
        // download waits for it; after an exception this wait does)
        if (exec_) (void)hipStreamSynchronize(s_);
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
    }
    BatchGraph(const BatchGraph&) = delete;
    BatchGraph& operator=(const BatchGraph&) = delete;

private:
    hipStream_t s_ = nullptr;
    bool capturing_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace

template <typename T>
int Workspace<T>::run_fast_batch(int count) {
    // Each level: r = A^T u - beta v_prev, DGKS against V, rank decision,
    // v = r / ||r|| appended to V; l = A v - alpha u, DGKS against u, rank
    // decision, u' = l / ||l||, g entry u'^T b (reference
    // krylov_workspace.cpp:589-831 for a one-column frontier).  A declined
    // decision clears the device flag `go`, which skips the rest of the
    // batch; the host commits the accepted prefix.
    hipStream_t s = ctx_.stream;
    Clock& clk = ctx_.clock;
    sync_projected_rhs();
    const int k0 = k_;
    ensure_basis(k0 + count);
    // partial buffers sized for the largest block counts of the batch
    int64_t nbmax = ceil_div(m_, chain_block_rows(m_, 1));
    int64_t qmax = nbmax;
    for (int j = 0; j < count; ++j) {
        const int64_t nbj = ceil_div(n_, chain_block_rows(n_, k0 + j));
        nbmax = std::max(nbmax, nbj);
        qmax = std::max<int64_t>(qmax, nbj * (k0 + j));
    }
    const int nbN = static_cast<int>(ceil_div(m_, chain_block_rows(m_, 1)));
    fast_cells_.ensure(static_cast<size_t>(4 + count));
    fast_rec_.ensure(static_cast<size_t>(8 * count));
    // statistics partials: one triple per block
    const int64_t spart = 3 * nbmax;
    fast_part_.ensure(static_cast<size_t>(4 * spart + 2 * qmax + static_cast<int64_t>(count) * nbN));
    double* cells = fast_cells_.get();
    double* go = cells, *alpha = cells + 1, *beta = cells + 2;  // beta[j]: coupling of level j
    double* sP = fast_part_.get();      // scale partials
    double* bP = sP + spart;            // before
    double* f1 = bP + spart;            // after the first pass
    double* f2 = f1 + spart;            // after the second pass
    double* Q1 = f2 + spart;            // projection partials, first pass
    double* Q2 = Q1 + qmax;             // second pass
    double* gpart = Q2 + qmax;          // g partials, nbN per level
    // Graph capture needs every product to be enqueued without host work on
    // the stream: no event pairs (operator profiling, event-timed products)
    // and stamp room reserved beforehand.
    const bool graph = ctx_.level_graphs && !op_.profiling() &&
                       (ctx_.clock.synthetic() || op_.times_own_kernels());
    if (graph && !ctx_.clock.synthetic()) ctx_.timer.reserve(s, 2 * static_cast<size_t>(count));
    BatchGraph batch_graph(graph ? s : nullptr);
    fill<double>(s, fast_rec_.get(), 8 * static_cast<int64_t>(count), 0.0);
    fill<double>(s, go, 1, 1.0);
    fill<double>(s, beta, 1, coupling_(0, 0));
    T* buf[2] = {left_.ptr(), work_.ptr()};
    const int64_t ldbuf[2] = {left_.ld(), work_.ld()};
    T* r = right_.ptr();
    const int64_t ldr = right_.ld();
    T* V = V_.ptr();
    const int64_t ldv = V_.ld();
    const double fac_l = rank_factor(1, 1);
    const double dgks = std::sqrt(0.5);
    for (int j = 0; j < count; ++j) {
        const int k = k0 + j;
        T* u = buf[j % 2];
        T* l = buf[(j + 1) % 2];
        const int64_t ldu = ldbuf[j % 2], ldl = ldbuf[(j + 1) % 2];
        double* rec = fast_rec_.get() + 8 * j;
        // right residual against V (n x k)
        enqueue_product(u, ldu, r, ldr, true);
        chain_couple_project<T>(s, n_, k, r, ldr, V + (k - 1), ldv, beta + j, V, ldv, go, sP, bP, Q1);
        chain_update<T>(s, n_, k, V, ldv, Q1, r, ldr, f1, Q2, go, nullptr, nullptr, 0.0);
        chain_update<T>(s, n_, k, V, ldv, Q2, r, ldr, f2, nullptr, go, bP, f1, dgks);
        chain_normalize<T>(s, n_, k, r, ldr, V + k, ldv, sP, bP, f1, f2, dgks, rank_factor(k, 1), nullptr, nullptr,
                           nullptr, nullptr, rec, alpha, go);
        // left residual against u
        enqueue_product(V + k, ldv, l, ldl, false);
        chain_couple_project<T>(s, m_, 1, l, ldl, u, ldu, alpha, u, ldu, go, sP, bP, Q1);
        chain_update<T>(s, m_, 1, u, ldu, Q1, l, ldl, f1, Q2, go, nullptr, nullptr, 0.0);
        chain_update<T>(s, m_, 1, u, ldu, Q2, l, ldl, f2, nullptr, go, bP, f1, dgks);
        chain_normalize<T>(s, m_, 1, l, ldl, l, ldl, sP, bP, f1, f2, dgks, fac_l, Q1, Q2, b_dev_,
                           gpart + static_cast<int64_t>(j) * nbN, rec + 4, beta + j + 1, go);
    }
    batch_graph.launch();
    fast_gpart_.resize(static_cast<size_t>(count) * nbN);
    fast_host_.resize(static_cast<size_t>(8 * count));
    AMFLS_HIP_CHECK(hipMemcpyAsync(fast_gpart_.data(), gpart, fast_gpart_.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
    ctx_.download(fast_host_.data(), fast_rec_.get(), fast_host_.size());
    // commit the accepted prefix exactly as advance_level would have
    int done = 0;
    for (; done < count; ++done) {
        const double* rec = fast_host_.data() + 8 * done;
        if (rec[2] != 1.0 || rec[6] != 1.0) break;
        const double alpha = rec[0], scale_r = rec[1], beta_new = rec[4], scale_l = rec[5], c_l = rec[7];
        double gv = 0.0;
        for (int q = 0; q < nbN; ++q) gv += fast_gpart_[static_cast<size_t>(done) * nbN + q];
        const int k = k_;
        const int old_left = static_cast<int>(g_.size());
        for (int side = 0; side < 2; ++side) {
            const double start = clk.now();
            clk.charge_operator(op_.relative_block_product_cost(1));
            if (clk.synthetic()) (side == 0 ? stats_.at_seconds : stats_.a_seconds) += clk.now() - start;
        }
        stats_.at_columns += 1;
        stats_.search_at_columns += 1;
        ++stats_.at_block_calls;
        ++stats_.search_at_block_calls;
        stats_.a_columns += 1;
        stats_.search_a_columns += 1;
        ++stats_.a_block_calls;
        ++stats_.search_a_block_calls;
        clk.charge_local(1, k);
        clk.charge_local(1, 1);
        StepAudit a;
        a.requested_injected_columns = pending_requested_;
        a.retained_injected_columns = pending_retained_;
        pending_requested_ = pending_retained_ = 0;
        a.left_input_width = 1;
        a.basis_rank_before = k;
        a.basis_rank_after = k + 1;
        a.retained_right_width = 1;
        a.right_rank_cutoff = rank_cutoff(scale_r, k, 1);
        a.retained_next_left_width = 1;
        a.left_rank_cutoff = rank_cutoff(scale_l, 1, 1);
        audit_ = a;
        std::vector<double> col(2);
        col[0] = alpha + c_l;
        col[1] = beta_new;
        T_.lo.push_back(active_first_);
        T_.val.push_back(std::move(col));
        T_.rows = old_left + 1;
        g_.push_back(gv);
        g_host_ = g_.size();
        k_ += 1;
        active_first_ = old_left;
        active_width_ = 1;
        prev_first_ = k;
        prev_width_ = 1;
        coupling_ = HostMat(1, 1);
        coupling_(0, 0) = beta_new;
        ++depth_;
    }
    if (done % 2 == 1) std::swap(left_, work_);
    return done;
}

template <typename T>
bool Workspace<T>::advance_toward(int target_depth) {
    const int want = target_depth - depth_;
    if (want <= 0) return advance_level();
    const int count = std::min(want, fast_batch_limit());
    if (count <= 0) return advance_level();
    const int done = run_fast_batch(count);
    if (done == count) return true;
    return advance_level();  // the declined level, host-driven
}

template <typename T>
void Workspace<T>::candidate(double* x_dev) {
    const double rcond = std::min(0.5, 64.0 * std::numeric_limits<double>::epsilon() *
                                           std::max(T_.rows, T_.cols()));
    if (k_ == 0) {
        fill<double>(ctx_.stream, x_dev, n_, 0.0);
        return;
    }
    bool dense = false;
    sync_projected_rhs();
    if (!projected_solve(T_, g_, lambda_, rcond, y_, &dense)) {
        y_.assign(k_, std::numeric_limits<double>::quiet_NaN());
    }
    double* yd = ctx_.upload(y_.data(), y_.size());
    gemm_nn<double, double>(ctx_.stream, n_, 1, k_, Strided{V_.ptr(), vt_of(T{}), V_.ld(), 1}, yd, 1, x_dev, 1,
                            1.0, 0.0);
}

template <typename T>
LeastSquaresResult Workspace<T>::evaluate(double tolerance, int epoch, int auxiliary_width) {
    if (!initialized_) throw std::logic_error("block Golub-Kahan workspace is not initialized");
    if (!std::isfinite(tolerance) || tolerance <= 0.0 || tolerance >= 1.0)
        throw std::invalid_argument("block Golub-Kahan tolerance must lie strictly between zero and one");
    ctx_.timer.flush();
    Clock& clk = ctx_.clock;
    const double start = clk.now();
    xcand_.ensure(static_cast<size_t>(n_));
    candidate(xcand_.get());
    if (!clk.synthetic()) ctx_.sync();
    stats_.projected_solve_seconds += clk.now() - start;
    LeastSquaresResult res = validator_.validate(xcand_.get(), lambda_, tolerance);
    res.depth = depth_;
    res.auxiliary_width = auxiliary_width;
    res.basis_rank = k_;
    res.trace.push_back(make_iteration_record(res, epoch, depth_, auxiliary_width));
    return res;
}

template class DeviceBlock<double>;
template class DeviceBlock<float>;
template class Workspace<double>;
template class Workspace<float>;

}  // namespace amflsx
