// Host-side launchers for the AMFLS gfx950 kernels.  Every launcher enqueues
// work on `stream` and returns without synchronizing; results that the host
// needs are copied out by the caller.
#pragma once

#include <vector>

#include "common.hpp"

namespace amflsx {

enum class ValueType : int { f64 = 0, f32 = 1, bf16 = 2 };

inline size_t value_size(ValueType t) {
    return t == ValueType::f64 ? 8 : (t == ValueType::f32 ? 4 : 2);
}

// Stride description of a large operand A(i, j) = base[i * rs + j * cs].
struct Strided {
    const void* ptr = nullptr;
    ValueType type = ValueType::f64;
    int64_t rs = 0;
    int64_t cs = 0;
    bool row_major() const { return cs == 1; }
};

// Device scratch owned by one engine instance (grown on demand).
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    // A fixed scratch never grows: ensure() throws std::length_error when
    // `bytes` is too small, and launchers that split work into chunks of
    // partial sums use no more chunks than fit (fit_chunks).
    bool fixed = false;
    void ensure(size_t need);
    void release();
};

// Number of chunks, each needing `per_chunk_bytes` of scratch, to launch:
// `want`, or fewer when a fixed scratch holds fewer (at least one fits, or
// std::length_error).
int64_t fit_chunks(const Scratch& scratch, size_t per_chunk_bytes, int64_t want);

// This device's allocation-free scratch (noalloc.hip): a view of a
// file-scope device array of kNoAllocScratchBytes, resolved with
// hipGetSymbolAddress once per device on first use; fixed, never freed.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
Scratch noalloc_scratch();

// ---------------------------------------------------------------- GEMMs
// Y(M x N, T row-major) = alpha * A(M x K) * B(K x N, TB row-major) + beta * Y.
template <typename T, typename TB>
void gemm_nn(hipStream_t s, int64_t M, int N, int64_t K, const Strided& A, const TB* B, int64_t ldb,
             T* Y, int64_t ldy, double alpha, double beta, const double* gate = nullptr);

// C(K x N) = A(M x K)^T * B(M x N, TB row-major); accumulation in double with a
// deterministic split-M reduction through `scratch`; C written as `ctype`
// (double or float) row-major with leading dimension ldc, C = alpha * sum.
// Launchers taking `gate` skip their work when *gate == 0 on the device.
template <typename TB>
void gemm_tn(hipStream_t s, int64_t M, int64_t K, int N, const Strided& A, const TB* B, int64_t ldb,
             void* C, ValueType ctype, int64_t ldc, double alpha, Scratch& scratch,
             const double* gate = nullptr);

// Tall-skinny products on the matrix cores (tall.hip) for row-major blocks
// with K, N <= 64: C = alpha A^T B (FP64 accumulation, partials through
// `scratch`) and Y = alpha A C + beta Y (C FP64 on the device, K x N with
// leading dimension ldc; Y may alias A).
bool tall_wide_supported(int K, int N);
template <typename T>
void tall_tn_wide(hipStream_t s, int64_t M, int K, int N, const T* A, int64_t rs, const T* B, int64_t ldb, void* C,
                  ValueType ctype, int64_t ldc, double alpha, Scratch& scratch, const double* gate = nullptr);
template <typename T>
void tall_nn_wide(hipStream_t s, int64_t M, int N, int K, const T* A, int64_t rs, const double* C, int64_t ldc, T* Y,
                  int64_t ldy, double alpha, double beta, const double* gate = nullptr);
// y = alpha V t + beta y and c = V^T g for a row-major basis V (n x k) and
// FP64 vectors (tall.hip).
template <typename TV>
void basis_apply(hipStream_t s, int64_t n, int k, const TV* V, int64_t ldv, const double* t, double* y, double alpha,
                 double beta);
template <typename TV>
void basis_project(hipStream_t s, int64_t n, int k, const TV* V, int64_t ldv, const double* g, double* c,
                   Scratch& scratch);

// In-place safe row-local product: Y[i, 0:r] = X[i, 0:w] * M (w x r, double
// row-major, device).  w, r <= 128.
template <typename T>
void rowmul(hipStream_t s, int64_t rows, int w, int r, const T* X, int64_t ldx, const double* M,
            T* Y, int64_t ldy);

// ---------------------------------------------------------------- dense, matrix cores
// Row-major A (m x n; bf16, fp32 or fp64 storage) times row-major blocks of
// at most 64 columns on MFMA (dense_mfma.hip): split3 = fp32-accurate
// products from bf16 A and three-term bf16 splits of the block; f32 = fp32
// matrix cores; f64 = fp64 matrix cores on A upconverted exactly.
enum class DenseMfmaMode : int { split3 = 0, f32 = 1, f64 = 2 };
// The mode for a storage type and block precision; `reduced` asks for
// fp32-accurate products on fp64 blocks.
// Narrow dense products on the vector ALUs (dense_narrow.hip): row-major A
// (bf16 / fp32 / fp64) and blocks of 1..8 columns, FP64 accumulation.
bool dense_narrow_supported(const Strided& A, int ncols);
template <typename TX>
void dense_narrow_nn(hipStream_t s, const Strided& A, int64_t M, int64_t N, int ncols, const TX* X, int64_t ldx,
                     TX* Y, int64_t ldy);
template <typename TX>
void dense_narrow_tn(hipStream_t s, const Strided& A, int64_t M, int64_t N, int ncols, const TX* Y, int64_t ldy,
                     TX* Z, int64_t ldz, Scratch& scratch);
DenseMfmaMode dense_mfma_mode(ValueType storage, bool fp64_blocks, bool reduced);
template <typename TX>
void dense_mfma_nn(hipStream_t s, DenseMfmaMode mode, const Strided& A, int64_t M, int64_t N, int ncols, const TX* X,
                   int64_t ldx, TX* Y, int64_t ldy, Scratch& scratch);
template <typename TX>
void dense_mfma_tn(hipStream_t s, DenseMfmaMode mode, const Strided& A, int64_t M, int64_t N, int ncols, const TX* Y,
                   int64_t ldy, TX* Z, int64_t ldz, Scratch& scratch);
// Upper bound of the scratch bytes that a product of an M x N row-major dense
// operator with a block of at most max_cols (<= 64) columns takes (partial
// sums of the A^T products, split planes of X); the operator grows its
// context's scratch to it once, so no product allocates (or frees, which
// waits for the device) in the middle of a timed solve.
size_t dense_product_scratch_bound(int64_t M, int64_t N, int max_cols);

// ---------------------------------------------------------------- vectors
template <typename Tin, typename Tout>
void copy_block(hipStream_t s, int64_t rows, int cols, const Tin* X, int64_t ldx, Tout* Y,
                int64_t ldy, double scale);
template <typename T>
void fill(hipStream_t s, T* y, int64_t n, T value);
// Y += a * X on row-major blocks.
template <typename T>
void axpy_block(hipStream_t s, int64_t rows, int cols, double a, const T* X, int64_t ldx, T* Y, int64_t ldy);
void sub_vec(hipStream_t s, int64_t n, const double* a, const double* b, double* r);
// y = fma(alpha, x, z) elementwise, fused (matches std::fma).
void fma_vec(hipStream_t s, int64_t n, double alpha, const double* x, const double* z, double* y);
void add_vec(hipStream_t s, int64_t n, const double* x, double* y, const double* gate = nullptr);
void scale_vec(hipStream_t s, int64_t n, double alpha, const double* x, double* y);

// Norm statistics of a block viewed as one vector (rows*cols values):
// out[0] = max |v| over finite values, out[1] = sum (|v| 2^-e)^2 with e the
// frexp exponent of out[0], out[2] = number of nonfinite values.  One
// launch; `counter` is a device word that holds 0 between launches.
template <typename T>
void vector_stats(hipStream_t s, int64_t rows, int cols, const T* X, int64_t ldx, double* out_dev,
                  Scratch& scratch, unsigned* counter, const double* gate = nullptr);
// *gate = 1 if both statistics triples give finite norms and
// ||after|| < factor * ||before||, else 0 (device-side DGKS criterion).
void second_pass_gate(hipStream_t s, const double* before, const double* after, double factor, double* gate);

// ---------------------------------------------------------------- rank-one steps
// One launch each; `counter` is a device word that holds 0 between launches
// and gates (a, b) skip the launch while either holds 0.
// c = B^T w (B row-major N x k), coeff = c or coeff += c.
template <typename TB, typename TW>
void rank_one_project(hipStream_t s, int64_t N, int k, const TB* B, int64_t ldb, const TW* w, int64_t ldw,
                      double* c, double* coeff, bool accumulate, Scratch& scratch, unsigned* counter,
                      const double* gate_a = nullptr, const double* gate_b = nullptr);
// w -= B c (one rounding to T), statistics of the new w into out[0..2]; with
// gate_out, *gate_out = (||w|| < factor * ||before||, both finite).
template <typename T>
void rank_one_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* c, T* w,
                     int64_t ldw, double* out, Scratch& scratch, unsigned* counter, const double* gate_a = nullptr,
                     const double* gate_b = nullptr, const double* before = nullptr, double factor = 0.0,
                     double* gate_out = nullptr);

// ---------------------------------------------------------------- level chain
// Batched rank-one level kernels (level_chain.hip).  All kernels of one side
// of a level use blocks of chain_block_rows(N, k) rows, write one partial
// per block (statistics triples; k projection sums) and combine their
// producer's partials in every block; a kernel returns at once while
// *go == 0.
int chain_block_rows(int64_t N, int k);
// y = r - (*coef) p in place on r; statistics partials of r and y;
// projection partials Q of B^T y.
template <typename T>
void chain_couple_project(hipStream_t s, int64_t N, int k, T* r, int64_t ldr, const T* p, int64_t ldp,
                          const double* coef, const T* B, int64_t ldb, const double* go, double* part_r,
                          double* part_y, double* Q);
// y -= B (sum_b Qin[b]); statistics partials of y; with Qout, projection
// partials of the new y.  With before/after given it is the DGKS second
// pass and runs only when ||after|| < factor ||before||.
template <typename T>
void chain_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* Qin, T* y, int64_t ldy,
                  double* part_y, double* Qout, const double* go, const double* before, const double* after,
                  double factor);
// Rank decision and normalization (dst = y / ||y||), see level_chain.hip.
template <typename T>
void chain_normalize(hipStream_t s, int64_t N, int k, const T* y, int64_t ldy, T* dst, int64_t ldd,
                     const double* part_scale, const double* part_before, const double* part_first,
                     const double* part_second, double dgks_factor, double rank_factor, const double* Q1,
                     const double* Q2, const double* b, double* gpart, double* record, double* cell, double* go);

// ---------------------------------------------------------------- width MLP
// 12 -> 64 -> 64 -> 1 ReLU network of the learned width controller
// (mlp.hip).  params: flat FP32 array of kMlpParams values; norm: per-input
// mean (12) then inverse standard deviation (12); X: n x 12 row-major.
constexpr int kMlpInputs = 12, kMlpHidden = 64;
constexpr int kMlpParams = kMlpHidden * kMlpInputs + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;
void mlp_forward(hipStream_t s, const float* params, const float* norm, int64_t n, const float* X, float* logits);
// grad = d(mean BCE)/d(params), *loss = mean BCE (device pointers).
void mlp_gradient(hipStream_t s, const float* params, const float* norm, int64_t n, const float* X, const float* y,
                  float* grad, float* loss, Scratch& scratch);
// One Adam step (step counts from 1).
void mlp_adam(hipStream_t s, float* params, float* m, float* v, const float* grad, float lr, int step);

// Counter-addressed Gaussian block, index = (first_col + j) * rows + i.
// Counter-addressed synthetic CSR (generate.hip): row pointers (rows + 1,
// device) for row lengths in [lo, hi] (hi <= 64), returning nnz (waits);
// then sorted distinct uniform columns and N(0,1) values times log-uniform
// row scales in [scale_lo, scale_hi].
int64_t csr_generate_rows(hipStream_t s, int64_t rows, int64_t cols, int lo, int hi, uint64_t seed,
                          int64_t* row_ptr, Scratch& scratch);
void csr_generate_entries(hipStream_t s, int64_t rows, int64_t cols, uint64_t seed, double scale_lo,
                          double scale_hi, const int64_t* row_ptr, int32_t* col_idx, void* values, ValueType type);

template <typename T>
void gaussian_block(hipStream_t s, int64_t rows, int cols, uint64_t seed, uint64_t stream,
                    uint64_t first_col, T* Y, int64_t ldy);

// ---------------------------------------------------------------- CSR
struct CsrView {
    int64_t rows = 0;
    int64_t cols = 0;
    int64_t nnz = 0;
    const int64_t* row_ptr = nullptr;
    const int32_t* col_idx = nullptr;
    const void* values = nullptr;
    ValueType type = ValueType::f32;
    // Row partition of the width-one stream kernel (stream_blocks + 1 row
    // boundaries, device), or null.
    const int64_t* stream_rows = nullptr;
    int64_t stream_blocks = 0;
    // the same for the wide-budget stream kernel (8192 entries a block)
    const int64_t* stream_rows_wide = nullptr;
    int64_t stream_blocks_wide = 0;
    // Slab-tiled copy of the entries for the width-one slab kernel, or null.
    // Rows are cut into panels (slab_panel_count + 1 row boundaries) and
    // columns into slab_count slabs of csr_slab_cols() columns; tile
    // (panel p, slab k) holds entries slab_tiles[p * slab_count + k] ..
    // slab_tiles[p * slab_count + k + 1] - 1, in (row, column) order, each
    // index packed as (row - first row of the panel) << 16 | (column - first
    // column of the slab), values in A's type.
    const int64_t* slab_panels = nullptr;
    const int64_t* slab_tiles = nullptr;
    const uint32_t* slab_idx = nullptr;
    const void* slab_val = nullptr;
    int64_t slab_panel_count = 0;
    int slab_count = 0;
};

// Row boundaries (host) of the stream kernel's partition: consecutive rows
// whose entries fit one block's budget (csr_stream_budget(), or 8192 for the
// wide kernel), at most 1024 (4096) rows per block; a longer row gets a
// block of its own.
std::vector<int64_t> csr_stream_partition(const std::vector<int64_t>& row_ptr, bool wide = false);
int csr_stream_budget();
// Index of the width-one stream variant (CSR-Stream: entries streamed into
// LDS with coalesced loads, then summed per row).
int csr_spmm_stream_variant();

// Width-one slab kernel (x staged in LDS one column slab at a time).
int csr_spmm_slab_variant();
int csr_slab_cols();  // columns per slab
// Slabs of a matrix with `cols` columns, or -1 beyond the format's limit.
int csr_slab_count(int64_t cols);
// Panel row boundaries (host) for the slab copy: consecutive rows with at
// most `panel_entries` entries (a longer row gets a panel of its own) and at
// most the kernel's row capacity.
std::vector<int64_t> csr_slab_partition(const std::vector<int64_t>& row_ptr, int64_t panel_entries);
// Entries per panel for a matrix of nnz entries.
int64_t csr_slab_panel_entries(int64_t nnz);
// Builds the slab copy of A on `s` from device panel boundaries: tiles
// (npanels * slab_count + 1 offsets), packed indices and values (nnz each,
// values of A's type).  Stable, so every tile keeps A's (row, column) order.
void csr_slab_build(hipStream_t s, const CsrView& A, const int64_t* panels, int64_t npanels, int slab_count,
                    int64_t* tiles, uint32_t* idx, void* val, Scratch& scratch);

// Wall-clock stamps of one product taken inside its kernels (device memory,
// zeroed when allocated).  The stamps of product i live at slots[i + j *
// kStampStride]: j = 0 is the device wall clock when block 0 of the first
// launch begins; j = 1..kStampEnds are end slots, and every block of the
// final launch raises end slot 1 + (block % kStampEnds) to its own end time
// with an atomic maximum, so the largest end slot is the last block to
// finish whatever the row lengths.  The end slots of one product lie
// kStampStride values apart (distinct cache lines and memory channels), so
// the atomics of a 3000-block grid spread over kStampEnds addresses instead
// of serialising on one.  The product's time is max(end slots) - slot 0, in
// ticks of hipDeviceAttributeWallClockRate.
constexpr int kStampEnds = 64;
constexpr int kStampSlots = 1 + kStampEnds;
constexpr int64_t kStampStride = 4096;  // products per stamp buffer
struct ProductStamp {
    unsigned long long* slots = nullptr;  // slot 0 of one product, or null: untimed
    bool first = true;                    // this launch writes slot 0
    bool last = true;                     // this launch writes the end slots
};

// Y(rows x ncols) = A * X (+ add_scale * Add[0:rows]); X, Y, Add row-major.
// With ev_start / ev_stop, the events take the start and end timestamps of
// the product's kernel dispatches themselves (hipExtLaunchKernel), so their
// elapsed time is kernel execution without queue gaps.  With `stamp` (the
// kStampSlots values of one product), the kernels stamp their start and end.
template <typename T>
void csr_spmm(hipStream_t s, const CsrView& A, int ncols, const T* X, int64_t ldx, T* Y, int64_t ldy,
              const T* Add, int64_t ld_add, double add_scale, int variant, hipEvent_t ev_start = nullptr,
              hipEvent_t ev_stop = nullptr, unsigned long long* stamp = nullptr);
int csr_spmm_variant_count();
bool csr_spmm_variant_supports(int variant, int ncols);
const char* csr_spmm_variant_name(int variant);
// Entry lanes per row (p), column lanes (lc) and columns per lane (cpl).
bool csr_spmm_variant_shape(int variant, int* p, int* lc, int* cpl);
// Variant used when none is set: lanes sized to the block width and the
// mean number of entries per row.
int csr_spmm_default_variant(int ncols, double mean_row_nnz);

// CSR of A^T, stable (each transposed row lists original rows in increasing
// order).  Outputs sized cols+1, nnz, nnz (values of A's type).
void csr_transpose(hipStream_t s, const CsrView& A, int64_t* t_row_ptr, int32_t* t_col_idx,
                   void* t_values, Scratch& scratch);

// Row statistics of a CSR matrix in double: |a| sums, square sums per row,
// plus misc[0] = max |a|, misc[1] = nonfinite count.
void csr_row_stats(hipStream_t s, const CsrView& A, double* out_sum, double* out_sq, double* misc,
                   Scratch& scratch);
// Structure checks: out[0] row_ptr monotone and spanning, out[1] column
// indices in range, out[2] no duplicate column inside a row (1 = ok).
void csr_validate(hipStream_t s, const CsrView& A, int* out_dev);

// Dense statistics: per-row and per-column |a| sums and square sums
// (double), misc[0] = max |a|, misc[1] = nonfinite count.
void dense_stats(hipStream_t s, int64_t M, int64_t N, const Strided& A, double* row_sum,
                 double* row_sq, double* col_sum, double* col_sq, double* misc, Scratch& scratch);

}  // namespace amflsx
