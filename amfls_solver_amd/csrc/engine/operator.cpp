// Device operators: construction scans, products and validation models.
#include "operator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>

#include "fp_bounds.hpp"

namespace amflsx {

ValidationModel norm_model_from_sums(const std::vector<double>& row_sum, const std::vector<double>& row_sq,
                                     const std::vector<long long>& row_ops,
                                     const std::vector<double>& col_sum, const std::vector<double>& col_sq,
                                     const std::vector<long long>& col_ops, double max_entry,
                                     long long steps_apply, long long steps_transpose) {
    // Same bound chain as the reference constructors
    // (sparse_csr_operator.cpp:90-129): max row / column |a| sums bound
    // ||A||_inf and ||A||_1, the column square sums bound ||A||_F, and row /
    // column two-norms bound ||A||_2 from below.
    double max_row_upper = 0.0, max_col_upper = 0.0, two_norm_lower = 0.0, fro_sq_upper = 0.0;
    for (size_t i = 0; i < row_sum.size(); ++i) {
        max_row_upper = std::max(max_row_upper, fp::pos_sum_upper(row_sum[i], row_ops[i]));
        two_norm_lower = std::max(two_norm_lower, fp::down_sqrt(fp::pos_sum_lower(row_sq[i], row_ops[i])));
    }
    for (size_t j = 0; j < col_sum.size(); ++j) {
        max_col_upper = std::max(max_col_upper, fp::pos_sum_upper(col_sum[j], col_ops[j]));
        two_norm_lower = std::max(two_norm_lower, fp::down_sqrt(fp::pos_sum_lower(col_sq[j], col_ops[j])));
        fro_sq_upper = fp::up_add(fro_sq_upper, fp::pos_fma_square_sum_upper(col_sq[j], col_ops[j]));
    }
    ValidationModel m;
    m.steps_apply = steps_apply;
    m.steps_transpose = steps_transpose;
    if (fp::environment_ok()) {
        const double one_inf = fp::up_mul(fp::up_sqrt(max_col_upper), fp::up_sqrt(max_row_upper));
        m.upper = std::min(one_inf, fp::up_sqrt(fro_sq_upper));
        m.lower = std::max(max_entry, two_norm_lower);
    }
    if (!std::isfinite(m.upper) || m.lower > m.upper) {
        m.upper = kInfD;
        m.lower = 0.0;
    }
    return m;
}

// ------------------------------------------------------------------ profiling
void DeviceOperator::set_profiling(bool on) {
    std::lock_guard<std::mutex> g(prof_mu_);
    profiling_ = on;
    if (!on) resolve_pending();
}

void DeviceOperator::prof_begin(hipStream_t s) const {
    if (!profiling_) return;
    std::lock_guard<std::mutex> g(prof_mu_);
    hipEvent_t e;
    if (pool_.empty()) AMFLS_HIP_CHECK(hipEventCreate(&e));
    else { e = pool_.back(); pool_.pop_back(); }
    AMFLS_HIP_CHECK(hipEventRecord(e, s));
    open_start_ = e;
}

void DeviceOperator::prof_end(hipStream_t s, int ncols, bool transpose, int eb) const {
    if (!profiling_) return;
    std::lock_guard<std::mutex> g(prof_mu_);
    if (open_start_ == nullptr) return;
    hipEvent_t e;
    if (pool_.empty()) AMFLS_HIP_CHECK(hipEventCreate(&e));
    else { e = pool_.back(); pool_.pop_back(); }
    AMFLS_HIP_CHECK(hipEventRecord(e, s));
    pending_.push_back({open_start_, e, product_bytes(ncols, transpose, eb), ncols, eb, transpose});
    open_start_ = nullptr;
}

void DeviceOperator::prof_events(hipEvent_t* start, hipEvent_t* stop) const {
    *start = *stop = nullptr;
    if (!profiling_) return;
    std::lock_guard<std::mutex> g(prof_mu_);
    for (hipEvent_t* e : {start, stop}) {
        if (pool_.empty()) AMFLS_HIP_CHECK(hipEventCreate(e));
        else { *e = pool_.back(); pool_.pop_back(); }
    }
}

void DeviceOperator::prof_commit(hipEvent_t start, hipEvent_t stop, int ncols, bool transpose, int eb) const {
    if (start == nullptr || stop == nullptr) return;
    std::lock_guard<std::mutex> g(prof_mu_);
    pending_.push_back({start, stop, product_bytes(ncols, transpose, eb), ncols, eb, transpose});
}

void DeviceOperator::resolve_pending() {
    for (ProfPair& p : pending_) {
        AMFLS_HIP_CHECK(hipEventSynchronize(p.stop));
        float ms = 0.0f;
        AMFLS_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        ++launches_;
        seconds_ += 1e-3 * ms;
        bytes_ += p.bytes;
        records_.push_back({1e-3 * ms, p.bytes, p.ncols, p.elem_bytes, p.transpose});
        pool_.push_back(p.start);
        pool_.push_back(p.stop);
    }
    pending_.clear();
}

void DeviceOperator::kernel_stats(long long* launches, double* seconds, double* bytes) {
    std::lock_guard<std::mutex> g(prof_mu_);
    resolve_pending();
    *launches = launches_;
    *seconds = seconds_;
    *bytes = bytes_;
    launches_ = 0;
    seconds_ = bytes_ = 0.0;
    records_.clear();
}

void DeviceOperator::kernel_stats_class(bool transpose, int min_cols, int max_cols, int elem_bytes,
                                        long long* launches, double* seconds, double* bytes) {
    std::lock_guard<std::mutex> g(prof_mu_);
    resolve_pending();
    *launches = 0;
    *seconds = *bytes = 0.0;
    for (const ProfRecord& r : records_) {
        if (r.transpose != transpose || r.ncols < min_cols || r.ncols > max_cols || r.elem_bytes != elem_bytes)
            continue;
        ++*launches;
        *seconds += r.seconds;
        *bytes += r.bytes;
    }
}

// ------------------------------------------------------------------ CSR
CsrOperator::CsrOperator(int64_t rows, int64_t cols, int64_t nnz, const int64_t* row_ptr,
                         const int32_t* col_idx, const void* values, ValueType type, Context& ctx) {
    if (rows <= 0 || cols <= 0 || nnz < 0 || row_ptr == nullptr)
        throw std::invalid_argument("CSR operator dimensions are invalid");
    if (nnz > 0 && (col_idx == nullptr || values == nullptr))
        throw std::invalid_argument("CSR operator arrays are missing");
    if (type != ValueType::f32 && type != ValueType::f64)
        throw std::invalid_argument("CSR values must be float32 or float64");
    if (cols > (int64_t(1) << 31) - 1) throw std::invalid_argument("CSR column count exceeds int32");
    rows_ = rows;
    cols_ = cols;
    nnz_ = nnz;
    a_ = CsrView{rows, cols, nnz, row_ptr, col_idx, values, type};
    hipStream_t s = ctx.stream;

    DevBuf<int> flags(3);
    csr_validate(s, a_, flags.get());
    int hflags[3] = {0, 0, 0};
    ctx.download(hflags, flags.get(), 3);
    if (!hflags[0]) throw std::invalid_argument("CSR row offsets do not span the values or decrease");
    if (!hflags[1]) throw std::invalid_argument("CSR column index is out of range");
    if (!hflags[2]) throw std::invalid_argument("CSR rows must not contain duplicate columns");

    t_rp_.reset(static_cast<size_t>(cols + 1));
    t_ci_.reset(static_cast<size_t>(std::max<int64_t>(nnz, 1)));
    t_val_.reset(static_cast<size_t>(std::max<int64_t>(nnz, 1)) * value_size(type));
    csr_transpose(s, a_, t_rp_.get(), t_ci_.get(), t_val_.get(), ctx.scratch2);
    at_ = CsrView{cols, rows, nnz, t_rp_.get(), t_ci_.get(), t_val_.get(), type};

    DevBuf<double> rsum(rows), rsq(rows), csum(cols), csq(cols), misc(4);
    csr_row_stats(s, a_, rsum.get(), rsq.get(), misc.get(), ctx.scratch);
    csr_row_stats(s, at_, csum.get(), csq.get(), misc.get() + 2, ctx.scratch);
    std::vector<double> hrs(rows), hrq(rows), hcs(cols), hcq(cols), hm(4);
    std::vector<int64_t> hrp(rows + 1), htrp(cols + 1);
    ctx.download(hrs.data(), rsum.get(), rows);
    ctx.download(hrq.data(), rsq.get(), rows);
    ctx.download(hcs.data(), csum.get(), cols);
    ctx.download(hcq.data(), csq.get(), cols);
    ctx.download(hm.data(), misc.get(), 4);
    ctx.download(hrp.data(), row_ptr, rows + 1);
    ctx.download(htrp.data(), t_rp_.get(), cols + 1);
    if (hm[1] > 0.0) throw std::invalid_argument("CSR values must be finite");
    // row partitions of the width-one stream kernels (both budgets), for A and A^T
    std::vector<int64_t> parts[4];
    for (int side = 0; side < 2; ++side) {
        for (int wide = 0; wide < 2; ++wide) {
            std::vector<int64_t>& br = parts[2 * side + wide];
            br = csr_stream_partition(side == 0 ? hrp : htrp, wide != 0);
            DevBuf<int64_t>& dst = side == 0 ? (wide ? stream_a_wide_ : stream_a_) : (wide ? stream_at_wide_ : stream_at_);
            dst.reset(br.size());
            AMFLS_HIP_CHECK(hipMemcpyAsync(dst.get(), br.data(), br.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                           s));
            CsrView& v = side == 0 ? a_ : at_;
            if (wide) {
                v.stream_rows_wide = dst.get();
                v.stream_blocks_wide = static_cast<int64_t>(br.size()) - 1;
            } else {
                v.stream_rows = dst.get();
                v.stream_blocks = static_cast<int64_t>(br.size()) - 1;
            }
        }
    }
    ctx.sync();  // the host partitions stay alive until their copies are done
    std::vector<long long> rops(rows), cops(cols);
    long long max_row = 0, max_col = 0;
    for (int64_t i = 0; i < rows; ++i) {
        rops[i] = hrp[i + 1] - hrp[i];
        max_row = std::max(max_row, rops[i]);
    }
    for (int64_t j = 0; j < cols; ++j) {
        cops[j] = htrp[j + 1] - htrp[j];
        max_col = std::max(max_col, cops[j]);
    }
    model_ = norm_model_from_sums(hrs, hrq, rops, hcs, hcq, cops, hm[0], 2 * max_row + 2, 2 * max_col + 2);
}

int CsrOperator::variant_for(int s, bool transpose) const {
    auto it = variants_.find({s, transpose});
    return it == variants_.end() ? -1 : it->second;
}

void CsrOperator::set_variant(int s, bool transpose, int variant) {
    if (s == 1 && variant == csr_spmm_slab_variant() && !ensure_slab(transpose, true))
        throw std::length_error("the slab kernel takes at most 1024 slabs of 8192 columns and fewer than 2^31 entries");
    variants_[{s, transpose}] = variant;
}

bool CsrOperator::ensure_slab(bool transpose, bool force) {
    CsrView& v = transpose ? at_ : a_;
    if (v.slab_panels != nullptr) return true;
    const int nslab = csr_slab_count(v.cols);
    if (nslab < 0 || v.rows <= 0 || v.nnz >= (int64_t(1) << 31)) return false;
    std::vector<int64_t> hrp(static_cast<size_t>(v.rows + 1));
    AMFLS_HIP_CHECK(hipMemcpy(hrp.data(), v.row_ptr, hrp.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    const std::vector<int64_t> br = csr_slab_partition(hrp, csr_slab_panel_entries(v.nnz));
    const int64_t np = static_cast<int64_t>(br.size()) - 1;
    if (np * nslab >= (int64_t(1) << 31) - 1) return false;
    if (!force) {
        const double staging = static_cast<double>(np) * static_cast<double>(v.cols) * 8.0;
        const double entries = static_cast<double>(v.nnz) * (4.0 + static_cast<double>(value_size(v.type)));
        if (staging > 8.0 * entries) return false;
    }
    SlabCopy& c = transpose ? slab_at_ : slab_a_;
    c.panels.reset(static_cast<size_t>(np + 1));
    c.tiles.reset(static_cast<size_t>(np * nslab + 1));
    c.idx.reset(static_cast<size_t>(std::max<int64_t>(v.nnz, 1)));
    c.val.reset(static_cast<size_t>(std::max<int64_t>(v.nnz, 1)) * value_size(v.type));
    // a private stream and scratch: this runs outside any engine call
    hipStream_t s = nullptr;
    AMFLS_HIP_CHECK(hipStreamCreate(&s));
    Scratch scratch;
    struct Cleanup {
        hipStream_t s;
        Scratch& scratch;
        ~Cleanup() {
            (void)hipStreamSynchronize(s);
            scratch.release();
            (void)hipStreamDestroy(s);
        }
    } cleanup{s, scratch};
    AMFLS_HIP_CHECK(hipMemcpyAsync(c.panels.get(), br.data(), br.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    csr_slab_build(s, v, c.panels.get(), np, nslab, c.tiles.get(), c.idx.get(), c.val.get(), scratch);
    AMFLS_HIP_CHECK(hipStreamSynchronize(s));
    v.slab_panels = c.panels.get();
    v.slab_tiles = c.tiles.get();
    v.slab_idx = c.idx.get();
    v.slab_val = c.val.get();
    v.slab_panel_count = np;
    v.slab_count = nslab;
    return true;
}

namespace {

template <typename T>
std::vector<int> tune_widths(Context& ctx, const CsrView& a, const CsrView& at, int64_t rows, int64_t cols,
                             const std::vector<int>& widths, int reps,
                             std::map<std::pair<int, bool>, int>& variants) {
    std::vector<int> chosen;
    std::vector<hipEvent_t> ev(2 * static_cast<size_t>(reps));
    for (hipEvent_t& e : ev) AMFLS_HIP_CHECK(hipEventCreate(&e));
    try {
        const int64_t big = std::max(rows, cols);
        for (int s : widths) {
            if (s < 1 || s > 256) throw std::invalid_argument("autotune widths must lie in [1, 256]");
            DevBuf<T> in(static_cast<size_t>(big) * s), out(static_cast<size_t>(big) * s);
            fill<T>(ctx.stream, in.get(), big * s, T(1));
            int lanes = 1;
            while (lanes < s) lanes <<= 1;
            for (int tr = 0; tr < 2; ++tr) {
                const CsrView& A = tr ? at : a;
                const double mean = A.rows > 0 ? static_cast<double>(A.nnz) / A.rows : 0.0;
                int best = csr_spmm_default_variant(s, mean);
                float best_ms = std::numeric_limits<float>::infinity();
                for (int v = 0; v < csr_spmm_variant_count(); ++v) {
                    int p = 0, lc = 0, cpl = 0;
                    csr_spmm_variant_shape(v, &p, &lc, &cpl);
                    if (!csr_spmm_variant_supports(v, s)) continue;
                    if (s <= 64 && (cpl != 1 || (lc != lanes && lc != 2 * lanes))) continue;
                    csr_spmm<T>(ctx.stream, A, s, in.get(), s, out.get(), s, nullptr, 0, 0.0, v);  // warm
                    for (int r = 0; r < reps; ++r)
                        csr_spmm<T>(ctx.stream, A, s, in.get(), s, out.get(), s, nullptr, 0, 0.0, v, ev[2 * r],
                                    ev[2 * r + 1]);
                    ctx.sync();
                    std::vector<float> ms(static_cast<size_t>(reps));
                    for (int r = 0; r < reps; ++r) AMFLS_HIP_CHECK(hipEventElapsedTime(&ms[r], ev[2 * r], ev[2 * r + 1]));
                    std::nth_element(ms.begin(), ms.begin() + reps / 2, ms.end());
                    if (ms[reps / 2] < best_ms) {
                        best_ms = ms[reps / 2];
                        best = v;
                    }
                }
                variants[{s, tr != 0}] = best;
                chosen.push_back(best);
            }
        }
    } catch (...) {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return chosen;
}

}  // namespace

std::vector<int> CsrOperator::autotune(Context& ctx, const std::vector<int>& widths, bool fp64_blocks, int reps) {
    if (reps < 1) throw std::invalid_argument("autotune needs at least one timed launch");
    if (std::find(widths.begin(), widths.end(), 1) != widths.end()) {
        ensure_slab(false, false);
        ensure_slab(true, false);
    }
    ctx.sync();
    return fp64_blocks ? tune_widths<double>(ctx, a_, at_, rows_, cols_, widths, reps, variants_)
                       : tune_widths<float>(ctx, a_, at_, rows_, cols_, widths, reps, variants_);
}

void CsrOperator::apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const {
    hipEvent_t e0, e1;
    prof_events(&e0, &e1);
    unsigned long long* st = take_stamp();
    csr_spmm<double>(s, a_, n, X, ldx, Y, ldy, nullptr, 0, 0.0, variant_for(n, false), e0, e1, st);
    prof_commit(e0, e1, n, false, 8);
}
void CsrOperator::apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const {
    hipEvent_t e0, e1;
    prof_events(&e0, &e1);
    unsigned long long* st = take_stamp();
    csr_spmm<double>(s, at_, n, Y, ldy, X, ldx, nullptr, 0, 0.0, variant_for(n, true), e0, e1, st);
    prof_commit(e0, e1, n, true, 8);
}
void CsrOperator::apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const {
    hipEvent_t e0, e1;
    prof_events(&e0, &e1);
    unsigned long long* st = take_stamp();
    csr_spmm<float>(s, a_, n, X, ldx, Y, ldy, nullptr, 0, 0.0, variant_for(n, false), e0, e1, st);
    prof_commit(e0, e1, n, false, 4);
}
void CsrOperator::apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const {
    hipEvent_t e0, e1;
    prof_events(&e0, &e1);
    unsigned long long* st = take_stamp();
    csr_spmm<float>(s, at_, n, Y, ldy, X, ldx, nullptr, 0, 0.0, variant_for(n, true), e0, e1, st);
    prof_commit(e0, e1, n, true, 4);
}

bool CsrOperator::product_noalloc(hipStream_t s, bool transpose, const void* in, int n, int64_t ldi, void* out,
                                  int64_t ldo, bool fp64, Scratch& /*scratch: the SpMM needs none*/) const {
    const CsrView& v = transpose ? at_ : a_;
    const int var = variant_for(n, transpose);
    if (fp64)
        csr_spmm<double>(s, v, n, static_cast<const double*>(in), ldi, static_cast<double*>(out), ldo, nullptr, 0, 0.0,
                         var, nullptr, nullptr, nullptr);
    else
        csr_spmm<float>(s, v, n, static_cast<const float*>(in), ldi, static_cast<float*>(out), ldo, nullptr, 0, 0.0,
                        var, nullptr, nullptr, nullptr);
    return true;
}

template <typename T>
void CsrOperator::apply_t_add(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx, const T* add,
                              int64_t ld_add, double a) const {
    csr_spmm<T>(s, at_, n, Y, ldy, X, ldx, add, ld_add, a, variant_for(n, true), nullptr, nullptr, take_stamp());
}
template void CsrOperator::apply_t_add<double>(hipStream_t, const double*, int, int64_t, double*, int64_t,
                                               const double*, int64_t, double) const;
template void CsrOperator::apply_t_add<float>(hipStream_t, const float*, int, int64_t, float*, int64_t,
                                              const float*, int64_t, double) const;

double CsrOperator::product_bytes(int s, bool transpose, int eb) const {
    // values + column indices + row pointers + gathered input rows + output
    const double vb = static_cast<double>(value_size(a_.type));
    const double out_rows = transpose ? static_cast<double>(cols_) : static_cast<double>(rows_);
    const double in_rows = transpose ? static_cast<double>(rows_) : static_cast<double>(cols_);
    return (vb + 4.0) * static_cast<double>(nnz_) + 8.0 * (out_rows + 1.0) +
           static_cast<double>(eb) * s * (in_rows + out_rows);
}

void CsrOperator::ensure_host_copy() const {
    std::call_once(host_once_, [this] {
        h_rp_.resize(rows_ + 1);
        h_ci_.resize(nnz_);
        h_val_.resize(nnz_);
        AMFLS_HIP_CHECK(hipMemcpy(h_rp_.data(), a_.row_ptr, (rows_ + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (nnz_ > 0) {
            AMFLS_HIP_CHECK(hipMemcpy(h_ci_.data(), a_.col_idx, nnz_ * sizeof(int32_t), hipMemcpyDeviceToHost));
            if (a_.type == ValueType::f64) {
                AMFLS_HIP_CHECK(hipMemcpy(h_val_.data(), a_.values, nnz_ * sizeof(double), hipMemcpyDeviceToHost));
            } else {
                std::vector<float> tmp(nnz_);
                AMFLS_HIP_CHECK(hipMemcpy(tmp.data(), a_.values, nnz_ * sizeof(float), hipMemcpyDeviceToHost));
                for (int64_t e = 0; e < nnz_; ++e) h_val_[e] = tmp[e];
            }
        }
    });
}

bool CsrOperator::refine_apply(const long double* x, long double* y) const {
    ensure_host_copy();
    for (int64_t r = 0; r < rows_; ++r) {
        long double sum = 0.0L;
        for (int64_t e = h_rp_[r]; e < h_rp_[r + 1]; ++e)
            sum += static_cast<long double>(h_val_[e]) * x[h_ci_[e]];
        if (!std::isfinite(sum)) return false;
        y[r] = sum;
    }
    return true;
}

bool CsrOperator::refine_apply_t(const long double* y, long double* x) const {
    ensure_host_copy();
    std::fill(x, x + cols_, 0.0L);
    for (int64_t r = 0; r < rows_; ++r) {
        const long double in = y[r];
        if (!std::isfinite(in)) return false;
        for (int64_t e = h_rp_[r]; e < h_rp_[r + 1]; ++e) x[h_ci_[e]] += static_cast<long double>(h_val_[e]) * in;
    }
    for (int64_t j = 0; j < cols_; ++j)
        if (!std::isfinite(x[j])) return false;
    return true;
}

// ------------------------------------------------------------------ dense
DenseOperator::DenseOperator(int64_t rows, int64_t cols, const void* values, ValueType type, bool row_major,
                             int64_t ld, Context& ctx)
    : ctx_(ctx) {
    if (rows <= 0 || cols <= 0 || values == nullptr) throw std::invalid_argument("dense operator input is invalid");
    if (row_major ? ld < cols : ld < rows) throw std::invalid_argument("dense leading dimension is too small");
    rows_ = rows;
    cols_ = cols;
    a_.ptr = values;
    a_.type = type;
    a_.rs = row_major ? ld : 1;
    a_.cs = row_major ? 1 : ld;
    DevBuf<double> rsum(rows), rsq(rows), csum(cols), csq(cols), misc(2);
    dense_stats(ctx.stream, rows, cols, a_, rsum.get(), rsq.get(), csum.get(), csq.get(), misc.get(), ctx.scratch);
    std::vector<double> hrs(rows), hrq(rows), hcs(cols), hcq(cols), hm(2);
    ctx.download(hrs.data(), rsum.get(), rows);
    ctx.download(hrq.data(), rsq.get(), rows);
    ctx.download(hcs.data(), csum.get(), cols);
    ctx.download(hcq.data(), csq.get(), cols);
    ctx.download(hm.data(), misc.get(), 2);
    std::vector<long long> rops(rows, cols), cops(cols, rows);
    model_ = norm_model_from_sums(hrs, hrq, rops, hcs, hcq, cops, hm[0], 2 * cols + 2, 2 * rows + 2);
    if (hm[1] > 0.0) {  // reference: a nonfinite entry disables certification
        model_.upper = kInfD;
        model_.lower = 0.0;
    }
    // Grow the context's scratch now to what the widest product needs: a
    // product that grows it frees the old arena (which waits for the device)
    // and allocates inside a timed interval, so the first wide products of a
    // process used to look slow to the width controller and steer the first
    // solve onto another trajectory than every later one.
    if (a_.cs == 1) ctx.scratch.ensure(dense_product_scratch_bound(rows, cols, 64));
}

void DenseOperator::set_variant(int s, bool transpose, int variant) {
    if (variant < -1 || variant > 1) throw std::invalid_argument("dense product variant must be -1, 0 or 1");
    if (variant < 0) variants_.erase({s, transpose});
    else variants_[{s, transpose}] = variant;
}

bool DenseOperator::use_narrow(int ncols, bool transpose) const {
    if (!dense_narrow_supported(a_, ncols)) return false;
    auto it = variants_.find({ncols, transpose});
    // Default: the vector-ALU kernels for single columns only.  They are
    // also faster per product at widths 2-8 (1.3-3.4x on the 1e6 x 2000
    // bf16 matrix), but then the width controller, which weighs measured
    // per-level seconds, keeps that problem at width two and certifies in
    // 2.5 s instead of 1.0 s at width sixteen.
    if (it == variants_.end()) return ncols == 1;
    return it->second == 0;
}

template <typename T>
void DenseOperator::do_apply(hipStream_t s, const T* X, int n, int64_t ldx, T* Y, int64_t ldy, Scratch& scratch,
                             bool profile) const {
    if (profile) prof_begin(s);
    if (use_narrow(n, false))
        dense_narrow_nn<T>(s, a_, rows_, cols_, n, X, ldx, Y, ldy);
    else if (use_mfma(n))
        dense_mfma_nn<T>(s, dense_mfma_mode(a_.type, sizeof(T) == 8, reduced_), a_, rows_, cols_, n, X, ldx, Y, ldy,
                         scratch);
    else
        gemm_nn<T, T>(s, rows_, n, cols_, a_, X, ldx, Y, ldy, 1.0, 0.0);
    if (profile) prof_end(s, n, false, static_cast<int>(sizeof(T)));
}

template <typename T>
void DenseOperator::do_apply_t(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx, Scratch& scratch,
                               bool profile) const {
    const ValueType ct = sizeof(T) == 8 ? ValueType::f64 : ValueType::f32;
    if (profile) prof_begin(s);
    if (use_narrow(n, true))
        dense_narrow_tn<T>(s, a_, rows_, cols_, n, Y, ldy, X, ldx, scratch);
    else if (use_mfma(n))
        dense_mfma_tn<T>(s, dense_mfma_mode(a_.type, sizeof(T) == 8, reduced_), a_, rows_, cols_, n, Y, ldy, X, ldx,
                         scratch);
    else
        gemm_tn<T>(s, rows_, cols_, n, a_, Y, ldy, X, ct, ldx, 1.0, scratch);
    if (profile) prof_end(s, n, true, static_cast<int>(sizeof(T)));
}

void DenseOperator::apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const {
    do_apply<double>(s, X, n, ldx, Y, ldy, ctx_.scratch, true);
}
void DenseOperator::apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const {
    do_apply_t<double>(s, Y, n, ldy, X, ldx, ctx_.scratch, true);
}
void DenseOperator::apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const {
    do_apply<float>(s, X, n, ldx, Y, ldy, ctx_.scratch, true);
}
void DenseOperator::apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const {
    do_apply_t<float>(s, Y, n, ldy, X, ldx, ctx_.scratch, true);
}

bool DenseOperator::product_noalloc(hipStream_t s, bool transpose, const void* in, int n, int64_t ldi, void* out,
                                    int64_t ldo, bool fp64, Scratch& scratch) const {
    if (fp64) {
        if (transpose)
            do_apply_t<double>(s, static_cast<const double*>(in), n, ldi, static_cast<double*>(out), ldo, scratch, false);
        else
            do_apply<double>(s, static_cast<const double*>(in), n, ldi, static_cast<double*>(out), ldo, scratch, false);
    } else {
        if (transpose)
            do_apply_t<float>(s, static_cast<const float*>(in), n, ldi, static_cast<float*>(out), ldo, scratch, false);
        else
            do_apply<float>(s, static_cast<const float*>(in), n, ldi, static_cast<float*>(out), ldo, scratch, false);
    }
    return true;
}

double DenseOperator::product_bytes(int s, bool transpose, int eb) const {
    (void)transpose;
    return static_cast<double>(value_size(a_.type)) * static_cast<double>(rows_) * static_cast<double>(cols_) +
           static_cast<double>(eb) * s * static_cast<double>(rows_ + cols_);
}

bool DenseOperator::supports_refinement() const {
    // A host long-double pass is affordable up to ~64M entries.
    return static_cast<double>(rows_) * static_cast<double>(cols_) <= 6.4e7;
}

bool DenseOperator::stream_rows(const std::function<bool(int64_t, int64_t, const double*)>& use) const {
    const size_t es = value_size(a_.type);
    const int64_t chunk = std::min<int64_t>(rows_, std::max<int64_t>(1, (int64_t(1) << 22) / cols_));
    std::vector<char> raw(static_cast<size_t>(chunk * cols_) * es);
    std::vector<double> val(static_cast<size_t>(chunk * cols_));
    const char* base = static_cast<const char*>(a_.ptr);
    const bool row_major = a_.cs == 1;
    for (int64_t i0 = 0; i0 < rows_; i0 += chunk) {
        const int64_t nr = std::min(chunk, rows_ - i0);
        // one 2-D copy per chunk: nr rows of pitch rs (row-major) or cols
        // column segments of nr entries and pitch cs (column-major)
        if (row_major)
            AMFLS_HIP_CHECK(hipMemcpy2D(raw.data(), cols_ * es, base + i0 * a_.rs * es, a_.rs * es, cols_ * es, nr,
                                        hipMemcpyDeviceToHost));
        else
            AMFLS_HIP_CHECK(hipMemcpy2D(raw.data(), nr * es, base + i0 * es, a_.cs * es, nr * es, cols_,
                                        hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nr; ++i)
            for (int64_t j = 0; j < cols_; ++j) {
                const char* e = raw.data() + (row_major ? i * cols_ + j : j * nr + i) * es;
                double v;
                if (a_.type == ValueType::f64) {
                    std::memcpy(&v, e, 8);
                } else if (a_.type == ValueType::f32) {
                    float f;
                    std::memcpy(&f, e, 4);
                    v = f;
                } else {  // bf16: the high half of an fp32
                    uint16_t h;
                    std::memcpy(&h, e, 2);
                    const uint32_t u = uint32_t(h) << 16;
                    float f;
                    std::memcpy(&f, &u, 4);
                    v = f;
                }
                val[i * cols_ + j] = v;
            }
        if (!use(i0, nr, val.data())) return false;
    }
    return true;
}

bool DenseOperator::refine_apply(const long double* x, long double* y) const {
    return stream_rows([&](int64_t i0, int64_t nr, const double* v) {
        for (int64_t i = 0; i < nr; ++i) {
            long double sum = 0.0L;
            const double* row = v + i * cols_;
            for (int64_t j = 0; j < cols_; ++j) sum += static_cast<long double>(row[j]) * x[j];
            if (!std::isfinite(sum)) return false;
            y[i0 + i] = sum;
        }
        return true;
    });
}

bool DenseOperator::refine_apply_t(const long double* y, long double* x) const {
    std::fill(x, x + cols_, 0.0L);
    const bool ok = stream_rows([&](int64_t i0, int64_t nr, const double* v) {
        for (int64_t i = 0; i < nr; ++i) {
            const long double in = y[i0 + i];
            if (!std::isfinite(in)) return false;
            const double* row = v + i * cols_;
            for (int64_t j = 0; j < cols_; ++j) x[j] += static_cast<long double>(row[j]) * in;
        }
        return true;
    });
    if (!ok) return false;
    for (int64_t j = 0; j < cols_; ++j)
        if (!std::isfinite(x[j])) return false;
    return true;
}

// ------------------------------------------------------------------ augmented
AugmentedOperator::AugmentedOperator(const DeviceOperator& base, double a, Context& /*ctx*/)
    : base_(base), a_(a) {
    rows_ = base.rows() + base.cols();
    cols_ = base.cols();
    const ValidationModel b = base.validation_model();
    model_.upper = fp::up_sqrt(fp::up_add(fp::up_mul(b.upper, b.upper), fp::up_mul(a, a)));
    model_.lower = std::max(b.lower, fp::down_mul(a, 1.0));
    model_.steps_apply = std::max<long long>(b.steps_apply, 2);
    model_.steps_transpose = b.steps_transpose + 2;
    if (!std::isfinite(model_.upper)) {
        model_.upper = kInfD;
        model_.lower = 0.0;
    }
}

template <typename T>
void AugmentedOperator::do_apply(hipStream_t s, const T* X, int n, int64_t ldx, T* Y, int64_t ldy) const {
    // a CSR base stamps its own kernel (the a X rows' copy is not timed)
    base_.stamp_next(take_stamp());
    base_.apply(s, X, n, ldx, Y, ldy);
    copy_block<T, T>(s, base_.cols(), n, X, ldx, Y + base_.rows() * ldy, ldy, a_);
}

template <typename T>
void AugmentedOperator::do_apply_t(hipStream_t s, const T* Y, int n, int64_t ldy, T* X, int64_t ldx) const {
    const T* tail = Y + base_.rows() * ldy;
    if (auto* csr = dynamic_cast<const CsrOperator*>(&base_)) {
        csr->stamp_next(take_stamp());
        csr->apply_t_add<T>(s, Y, n, ldy, X, ldx, tail, ldy, a_);
        return;
    }
    base_.apply_t(s, Y, n, ldy, X, ldx);
    axpy_block<T>(s, base_.cols(), n, a_, tail, ldy, X, ldx);
}

void AugmentedOperator::apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const {
    do_apply<double>(s, X, n, ldx, Y, ldy);
}
void AugmentedOperator::apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const {
    do_apply_t<double>(s, Y, n, ldy, X, ldx);
}
void AugmentedOperator::apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const {
    do_apply<float>(s, X, n, ldx, Y, ldy);
}
void AugmentedOperator::apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const {
    do_apply_t<float>(s, Y, n, ldy, X, ldx);
}

}  // namespace amflsx

namespace amflsx {

HostOperator::HostOperator(int64_t rows, int64_t cols, HostApplyFn apply, HostApplyFn apply_transpose, void* user,
                           ValidationModel model, HostHooks hooks)
    : apply_(apply), apply_t_(apply_transpose), user_(user), model_(model), hooks_(hooks) {
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("operator dimensions must be positive");
    if (apply == nullptr || apply_transpose == nullptr) throw std::invalid_argument("operator callbacks are missing");
    if (std::isnan(model.upper) || std::isnan(model.lower) || model.lower < 0.0 || model.steps_apply < 0 ||
        model.steps_transpose < 0)
        throw std::invalid_argument("validation error model is invalid");
    rows_ = rows;
    cols_ = cols;
}

template <typename T>
void HostOperator::call(hipStream_t s, HostApplyFn fn, const T* in, int64_t in_rows, int n, int64_t ldi, T* out,
                        int64_t out_rows, int64_t ldo) const {
    if (n <= 0) return;
    const size_t in_bytes = static_cast<size_t>(in_rows) * n * sizeof(T);
    const size_t out_bytes = static_cast<size_t>(out_rows) * n * sizeof(T);
    hin_.ensure(in_bytes);
    hout_.ensure(out_bytes);
    T* hi = reinterpret_cast<T*>(hin_.get());
    T* ho = reinterpret_cast<T*>(hout_.get());
    // device row-major (ld ldi) -> host row-major packed
    AMFLS_HIP_CHECK(hipMemcpy2DAsync(hi, n * sizeof(T), in, ldi * sizeof(T), n * sizeof(T), in_rows,
                                     hipMemcpyDeviceToHost, s));
    AMFLS_HIP_CHECK(hipStreamSynchronize(s));
    cin_.resize(static_cast<size_t>(in_rows) * n);
    cout_.assign(static_cast<size_t>(out_rows) * n, 0.0);
    for (int64_t i = 0; i < in_rows; ++i)
        for (int j = 0; j < n; ++j) cin_[static_cast<size_t>(j) * in_rows + i] = static_cast<double>(hi[i * n + j]);
    if (fn(user_, cin_.data(), n, cout_.data()) != 0) throw std::runtime_error("host operator callback failed");
    for (int64_t i = 0; i < out_rows; ++i)
        for (int j = 0; j < n; ++j) ho[i * n + j] = static_cast<T>(cout_[static_cast<size_t>(j) * out_rows + i]);
    AMFLS_HIP_CHECK(hipMemcpy2DAsync(out, ldo * sizeof(T), ho, n * sizeof(T), n * sizeof(T), out_rows,
                                     hipMemcpyHostToDevice, s));
    // the page-locked buffer is reused by the next product: wait for the copy
    AMFLS_HIP_CHECK(hipStreamSynchronize(s));
}

void HostOperator::apply(hipStream_t s, const double* X, int n, int64_t ldx, double* Y, int64_t ldy) const {
    prof_begin(s);
    call<double>(s, apply_, X, cols_, n, ldx, Y, rows_, ldy);
    prof_end(s, n, false, 8);
}
void HostOperator::apply_t(hipStream_t s, const double* Y, int n, int64_t ldy, double* X, int64_t ldx) const {
    prof_begin(s);
    call<double>(s, apply_t_, Y, rows_, n, ldy, X, cols_, ldx);
    prof_end(s, n, true, 8);
}
void HostOperator::apply(hipStream_t s, const float* X, int n, int64_t ldx, float* Y, int64_t ldy) const {
    prof_begin(s);
    call<float>(s, apply_, X, cols_, n, ldx, Y, rows_, ldy);
    prof_end(s, n, false, 4);
}
void HostOperator::apply_t(hipStream_t s, const float* Y, int n, int64_t ldy, float* X, int64_t ldx) const {
    prof_begin(s);
    call<float>(s, apply_t_, Y, rows_, n, ldy, X, cols_, ldx);
    prof_end(s, n, true, 4);
}

}  // namespace amflsx
