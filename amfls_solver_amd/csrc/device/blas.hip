// Tall-skinny block kernels for the block Golub-Kahan engine on gfx950:
// the orthogonalization GEMMs (C = A^T B, Y -= A C, row-local R^{-1}
// products), dense-operator products, vector reductions for the certified
// norm intervals, and the counter-addressed Gaussian generator.
//
// Everything here is HBM-bound at the widths AMFLS uses (s <= 64): each
// kernel streams its tall operand once, keeps the small operand in LDS, and
// accumulates reductions in FP64 with a deterministic two-stage split.
#include "kernels.hpp"

#include <string>

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace amflsx {

void Scratch::ensure(size_t need) {
    if (need <= bytes) return;
    if (fixed)
        throw std::length_error("the product needs " + std::to_string(need) + " bytes of scratch; the fixed scratch holds " +
                                std::to_string(bytes));
    if (ptr) AMFLS_HIP_CHECK(hipFree(ptr));
    ptr = nullptr;
    size_t grow = std::max(need, bytes * 2);
    AMFLS_HIP_CHECK(hipMalloc(&ptr, grow));
    bytes = grow;
}

int64_t fit_chunks(const Scratch& scratch, size_t per_chunk_bytes, int64_t want) {
    if (!scratch.fixed) return want;
    const int64_t room = static_cast<int64_t>(scratch.bytes / std::max<size_t>(per_chunk_bytes, 1));
    if (room < 1)
        throw std::length_error("one chunk of partial sums needs " + std::to_string(per_chunk_bytes) +
                                " bytes; the fixed scratch holds " + std::to_string(scratch.bytes));
    return std::min(want, room);
}

void Scratch::release() {
    if (fixed) return;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

namespace {

constexpr int kThreads = 256;

template <typename V>
__device__ __forceinline__ double ld_d(const V* p, int64_t i) { return to_double(p[i]); }

template <typename Acc, typename V>
__device__ __forceinline__ Acc ld_acc(const V* p, int64_t i) { return load_as<Acc, V>(p[i]); }

template <typename T> struct AccOf { using type = T; };

// Optional device predicate: a kernel given a gate returns at once when the
// gate holds 0 (work skipped by a decision taken on the device).
__device__ __forceinline__ bool gated_off(const double* gate) { return gate != nullptr && *gate == 0.0; }

// ------------------------------------------------------- tall FP64 on MFMA
// FP64 tall blocks with K, N <= 32 on v_mfma_f64_16x16x4_f64.  Operand
// layout (wave64): lane l supplies row l % 16 of the A operand and column
// l % 16 of the B operand at reduction index q = l / 16; its four results are
// column l % 16, rows 4 r + q (r < 4).
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Y = alpha A C + beta Y (A: M x K, row stride rs; C: K x N, ldc; Y: M x N,
// ldy; K <= KM, N <= 16 NT): a wavefront per 16-row tile, TPW tiles per
// wavefront.  With KL = KM / 4, lane (i, q) loads A[row i][KL q .. KL q + KL)
// (contiguous; the four lane groups cover the row with full cache lines), so
// instruction t contracts inner index KL q + t; C[KL q + t][16 n + i] stays in
// registers.  Every load of a tile precedes its stores, so with beta = 0 Y
// may alias A (same rows, same stride).
template <int KM, int NT, int TPW>
__global__ __launch_bounds__(kThreads) void tall_nn_f64_kernel(int64_t M, int N, int K, const double* A, int64_t rs,
                                                               const double* __restrict__ C, int64_t ldc, double* Y,
                                                               int64_t ldy, double alpha, double beta,
                                                               const double* gate) {
    if (gated_off(gate)) return;
    constexpr int KL = KM / 4;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int i = lane % 16, q = lane / 16;
    double b[KL][NT];
#pragma unroll
    for (int t = 0; t < KL; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n)
            b[t][n] = (KL * q + t < K && 16 * n + i < N) ? C[(KL * q + t) * ldc + 16 * n + i] : 0.0;
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + wave) * TPW;
    double a[TPW][KL];
#pragma unroll
    for (int p = 0; p < TPW; ++p) {
        const int64_t row = (tile0 + p) * 16 + i;
        const double* src = A + row * rs + KL * q;
        if (row < M && KL * q + KL <= K && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
#pragma unroll
            for (int h = 0; h < KL / 2; ++h) {
                const double2 v = reinterpret_cast<const double2*>(src)[h];
                a[p][2 * h] = v.x;
                a[p][2 * h + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int t = 0; t < KL; ++t) a[p][t] = (row < M && KL * q + t < K) ? src[t] : 0.0;
        }
    }
#pragma unroll
    for (int p = 0; p < TPW; ++p) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < KL; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p][t], b[t][n], acc, 0, 0, 0);
            const int col = 16 * n + i;
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = (tile0 + p) * 16 + 4 * r + q;
                if (row >= M) continue;
                double v = alpha * acc[r];
                if (beta != 0.0) v = fma(beta, Y[row * ldy + col], v);
                Y[row * ldy + col] = v;
            }
        }
    }
}

// P[block][kk][j] = sum over the block's rows of A[row][kk] B[row][j] (A: M x
// K, row stride rs; B: M x N, ldb; K <= 16 KT, N <= 16 NT).  Lane (c, q) loads
// A[row][16 m + c] and B[row][16 n + c] of row 4u + q of its wavefront's
// 4U-row group (each group of 16 lanes reads contiguous bytes of one row), U
// groups in flight per trip; the four wavefronts' sums are added in LDS in
// wavefront order.
template <int KT, int NT, int U>
__global__ __launch_bounds__(kThreads) void tall_tn_f64_kernel(int64_t M, int K, int N, const double* __restrict__ A,
                                                               int64_t rs, const double* __restrict__ B, int64_t ldb,
                                                               double* __restrict__ P, int64_t rows_per_block,
                                                               const double* gate) {
    if (gated_off(gate)) return;
    constexpr int W = kThreads / kWave;
    constexpr int TILES = KT * NT;
    __shared__ double red[W][TILES * 256];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int c = lane % 16, q = lane / 16;
    const int64_t r_begin = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r_end = std::min<int64_t>(M, r_begin + rows_per_block);
    f64x4 acc[KT][NT];
#pragma unroll
    for (int m = 0; m < KT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int64_t r0 = r_begin + wave * 4 * U; r0 < r_end; r0 += W * 4 * U) {
        double a[U][KT], b[U][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = r0 + 4 * u + q;
#pragma unroll
            for (int m = 0; m < KT; ++m) a[u][m] = (row < r_end && 16 * m + c < K) ? A[row * rs + 16 * m + c] : 0.0;
#pragma unroll
            for (int n = 0; n < NT; ++n) b[u][n] = (row < r_end && 16 * n + c < N) ? B[row * ldb + 16 * n + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int m = 0; m < KT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][m], b[u][n], acc[m][n], 0, 0, 0);
    }
    // tile (m, n) holds rows kk = 16 m + 4 r + q, columns j = 16 n + c
#pragma unroll
    for (int m = 0; m < KT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][(m * NT + n) * 256 + (4 * r + q) * 16 + c] = acc[m][n][r];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < TILES; ++e) {
        const int o = e * 256 + threadIdx.x;
        const int tile = o / 256, m = tile / NT, n = tile % NT;
        const int kk = 16 * m + (o % 256) / 16, j = 16 * n + o % 16;
        if (kk < K && j < N) {
            double v = red[0][o];
#pragma unroll
            for (int w = 1; w < W; ++w) v += red[w][o];
            P[static_cast<int64_t>(blockIdx.x) * K * N + kk * N + j] = v;
        }
    }
}

// Widest K and N the engine sends to the tall kernels.  The kernels take
// K, N <= 32, but with 32-column local products this fast the reference width
// controller widens the dense 1e6 x 2000 problem to 64 columns late in the
// solve, where the step's rank decision truncates and the solve ends at its
// precision limit; with 16 it keeps at most 32 columns and certifies.
constexpr int kTallMax = 16;

void tall_nn_f64(hipStream_t s, int64_t M, int N, int K, const double* A, int64_t rs, const double* C, int64_t ldc,
                 double* Y, int64_t ldy, double alpha, double beta, const double* gate) {
    constexpr int TPW = 2;
    const int64_t waves = ceil_div(ceil_div(M, 16), TPW);
    const dim3 grid(static_cast<unsigned>(ceil_div(waves, kThreads / kWave)));
#define AMFLS_TALL_NN(KM, NT)                                                                                       \
    hipLaunchKernelGGL((tall_nn_f64_kernel<KM, NT, TPW>), grid, dim3(kThreads), 0, s, M, N, K, A, rs, C, ldc, Y, ldy, \
                       alpha, beta, gate)
    if (K <= 16) {
        if (N <= 16) AMFLS_TALL_NN(16, 1);
        else AMFLS_TALL_NN(16, 2);
    } else {
        if (N <= 16) AMFLS_TALL_NN(32, 1);
        else AMFLS_TALL_NN(32, 2);
    }
#undef AMFLS_TALL_NN
    AMFLS_LAUNCH_CHECK();
}

// Blocks and rows per block of tall_tn_f64_kernel: about 1024 blocks, each
// row range a whole number of the block's 4 x 4U-row trips.
constexpr int kTallTnU = 8;
inline void tall_tn_shape(int64_t M, int64_t& nb, int64_t& rows_per_block) {
    constexpr int64_t trip = (kThreads / kWave) * 4 * kTallTnU;
    nb = std::max<int64_t>(1, std::min<int64_t>(1024, ceil_div(std::max<int64_t>(M, 1), 64)));
    rows_per_block = ceil_div(ceil_div(std::max<int64_t>(M, 1), nb), trip) * trip;
    nb = ceil_div(std::max<int64_t>(M, 1), rows_per_block);
}

void tall_tn_f64(hipStream_t s, int64_t M, int K, int N, const double* A, int64_t rs, const double* B, int64_t ldb,
                 double* P, int64_t nb, int64_t rows_per_block, const double* gate) {
#define AMFLS_TALL_TN(KT, NT)                                                                                    \
    hipLaunchKernelGGL((tall_tn_f64_kernel<KT, NT, kTallTnU>), dim3(static_cast<unsigned>(nb)), dim3(kThreads), 0, \
                       s, M, K, N, A, rs, B, ldb, P, rows_per_block, gate)
    if (K <= 16) {
        if (N <= 16) AMFLS_TALL_TN(1, 1);
        else AMFLS_TALL_TN(1, 2);
    } else {
        if (N <= 16) AMFLS_TALL_TN(2, 1);
        else AMFLS_TALL_TN(2, 2);
    }
#undef AMFLS_TALL_TN
    AMFLS_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ gemm_nn
// Wide tile: 64 rows x 64 columns per 256-thread block, 4x4 per thread.
template <typename V, typename TB, typename T, typename Acc, bool AROW>
__global__ __launch_bounds__(kThreads) void gemm_nn_tile_kernel(
    int64_t M, int N, int64_t K, const V* __restrict__ A, int64_t rs, int64_t cs,
    const TB* __restrict__ B, int64_t ldb, T* __restrict__ Y, int64_t ldy, Acc alpha, Acc beta,
    const double* gate) {
    if (gated_off(gate)) return;
    constexpr int BM = 64, BN = 64, BK = 32;
    __shared__ Acc As[BK][BM + 1];
    __shared__ Acc Bs[BK][BN];
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * BM;
    const int col0 = blockIdx.y * BN;
    const int t = threadIdx.x;
    const int tx = t % 16, ty = t / 16;
    Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc(0);
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int l = 0; l < (BM * BK) / kThreads; ++l) {
            const int idx = t + l * kThreads;
            int r, kk;
            if (AROW) { kk = idx % BK; r = idx / BK; } else { r = idx % BM; kk = idx / BM; }
            const int64_t gr = row0 + r, gk = k0 + kk;
            Acc v = Acc(0);
            if (gr < M && gk < K) v = ld_acc<Acc>(A, gr * rs + gk * cs);
            As[kk][r] = v;
        }
#pragma unroll
        for (int l = 0; l < (BN * BK) / kThreads; ++l) {
            const int idx = t + l * kThreads;
            const int c = idx % BN, kk = idx / BN;
            const int64_t gk = k0 + kk;
            const int gc = col0 + c;
            Bs[kk][c] = (gk < K && gc < N) ? static_cast<Acc>(B[gk * ldb + gc]) : Acc(0);
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < BK; ++kk) {
            Acc a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t gr = row0 + ty * 4 + i;
        if (gr >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gc = col0 + tx * 4 + j;
            if (gc >= N) continue;
            T* out = Y + gr * ldy + gc;
            Acc v = alpha * acc[i][j];
            if (beta != Acc(0)) v = fma(beta, static_cast<Acc>(*out), v);
            *out = static_cast<T>(v);
        }
    }
}

// Tall-skinny product for narrow outputs (N <= 16): each block owns 64 rows;
// A is staged through LDS in 64-wide K tiles with loads coalesced along the
// contiguous dimension of A (k for row-major, rows for column-major); four
// threads share a row, split the K tile and combine with two xor-shuffles.
template <typename V, typename TB, typename T, typename Acc, int NMAX, bool AROW>
__global__ __launch_bounds__(kThreads) void ts_gemm_nn_kernel(
    int64_t M, int N, int64_t K, const V* __restrict__ A, int64_t rs, int64_t cs, const TB* __restrict__ B,
    int64_t ldb, T* __restrict__ Y, int64_t ldy, Acc alpha, Acc beta, const double* gate) {
    if (gated_off(gate)) return;
    constexpr int BM = 64, KT = 64;
    __shared__ Acc As[BM][KT + 1];
    __shared__ Acc Bs[KT][NMAX];
    const int t = threadIdx.x;
    const int r = t / 4, q = t % 4;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * BM;
    Acc acc[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) acc[j] = Acc(0);
    for (int64_t k0 = 0; k0 < K; k0 += KT) {
#pragma unroll
        for (int l = 0; l < (BM * KT) / kThreads; ++l) {
            const int idx = t + l * kThreads;
            int rr, kk;
            if (AROW) { kk = idx % KT; rr = idx / KT; } else { rr = idx % BM; kk = idx / BM; }
            const int64_t gr = row0 + rr, gk = k0 + kk;
            As[rr][kk] = (gr < M && gk < K) ? ld_acc<Acc>(A, gr * rs + gk * cs) : Acc(0);
        }
        for (int idx = t; idx < KT * NMAX; idx += kThreads) {
            const int kk = idx / NMAX, j = idx % NMAX;
            const int64_t gk = k0 + kk;
            Bs[kk][j] = (gk < K && j < N) ? static_cast<Acc>(B[gk * ldb + j]) : Acc(0);
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = q * (KT / 4); kk < (q + 1) * (KT / 4); ++kk) {
            const Acc a = As[r][kk];
#pragma unroll
            for (int j = 0; j < NMAX; ++j) acc[j] = fma(a, Bs[kk][j], acc[j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
        acc[j] += __shfl_xor(acc[j], 1);
        acc[j] += __shfl_xor(acc[j], 2);
    }
    const int64_t gr = row0 + r;
    if (gr >= M) return;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
        if (j % 4 != q || j >= N) continue;
        T* out = Y + gr * ldy + j;
        Acc v = alpha * acc[j];
        if (beta != Acc(0)) v = fma(beta, static_cast<Acc>(*out), v);
        *out = static_cast<T>(v);
    }
}

// Thin product for row-major A with K <= 16 and N <= 16 (block updates of
// tall frontiers, R^-1 applications): a thread per row reads its K
// contiguous entries, B sits in LDS, the N outputs are written
// contiguously; the k terms are summed in order.
template <typename V, typename TB, typename T, typename Acc, int KM, int NM>
__global__ __launch_bounds__(kThreads) void thin_nn_kernel(int64_t M, int N, int K, const V* __restrict__ A,
                                                           int64_t rs, const TB* __restrict__ B, int64_t ldb,
                                                           T* __restrict__ Y, int64_t ldy, Acc alpha, Acc beta,
                                                           const double* gate) {
    if (gated_off(gate)) return;
    __shared__ Acc Bs[KM][NM];
    for (int idx = threadIdx.x; idx < KM * NM; idx += kThreads) {
        const int kk = idx / NM, j = idx % NM;
        Bs[kk][j] = (kk < K && j < N) ? static_cast<Acc>(B[kk * ldb + j]) : Acc(0);
    }
    __syncthreads();
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (row >= M) return;
    Acc a[KM];
#pragma unroll
    for (int kk = 0; kk < KM; ++kk) a[kk] = kk < K ? ld_acc<Acc>(A, row * rs + kk) : Acc(0);
    Acc acc[NM];
#pragma unroll
    for (int j = 0; j < NM; ++j) acc[j] = Acc(0);
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
#pragma unroll
        for (int j = 0; j < NM; ++j) acc[j] = fma(a[kk], Bs[kk][j], acc[j]);
    T* out = Y + row * ldy;
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        if (j >= N) break;
        Acc v = alpha * acc[j];
        if (beta != Acc(0)) v = fma(beta, static_cast<Acc>(out[j]), v);
        out[j] = static_cast<T>(v);
    }
}

template <typename V, typename TB, typename T>
void gemm_nn_dispatch(hipStream_t s, int64_t M, int N, int64_t K, const V* A, int64_t rs,
                      int64_t cs, const TB* B, int64_t ldb, T* Y, int64_t ldy, double alpha,
                      double beta, const double* gate) {
    using Acc = T;
    if (M <= 0 || N <= 0) return;
    const bool row_major = (cs == 1);
    if constexpr (std::is_same_v<V, float> && std::is_same_v<T, float> && std::is_same_v<TB, double>) {
        // FP32 blocks times FP64 coefficients: the matrix-core tall kernel
        if (row_major && K <= (int64_t{1} << 30) && tall_wide_supported(static_cast<int>(K), N)) {
            tall_nn_wide<float>(s, M, N, static_cast<int>(K), A, rs, B, ldb, Y, ldy, alpha, beta, gate);
            return;
        }
    }
    if constexpr (std::is_same_v<V, double> && std::is_same_v<TB, double> && std::is_same_v<T, double>) {
        if (row_major && N <= kTallMax && K <= kTallMax) {
            tall_nn_f64(s, M, N, static_cast<int>(K), A, rs, B, ldb, Y, ldy, alpha, beta, gate);
            return;
        }
    }
    if (row_major && N <= 16 && K <= 16) {
        const dim3 grid(static_cast<unsigned>(ceil_div(M, kThreads)));
#define AMFLS_THIN(KM, NM)                                                                                  \
        hipLaunchKernelGGL((thin_nn_kernel<V, TB, T, Acc, KM, NM>), grid, dim3(kThreads), 0, s, M, N,            \
                           static_cast<int>(K), A, rs, B, ldb, Y, ldy, Acc(alpha), Acc(beta), gate)
        const int kb = K <= 1 ? 1 : (K <= 2 ? 2 : (K <= 4 ? 4 : (K <= 8 ? 8 : 16)));
        const int nb = N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : (N <= 8 ? 8 : 16)));
        switch (kb * 100 + nb) {
            case 101: AMFLS_THIN(1, 1); break;
            case 202: AMFLS_THIN(2, 2); break;
            case 404: AMFLS_THIN(4, 4); break;
            case 808: AMFLS_THIN(8, 8); break;
            case 1616: AMFLS_THIN(16, 16); break;
            default:
                if (kb <= 4 && nb <= 4) AMFLS_THIN(4, 4);
                else if (kb <= 8 && nb <= 8) AMFLS_THIN(8, 8);
                else AMFLS_THIN(16, 16);
        }
#undef AMFLS_THIN
        AMFLS_LAUNCH_CHECK();
        return;
    }
    if (N <= 16) {
        const dim3 grid(static_cast<unsigned>(ceil_div(M, 64)));
#define AMFLS_TSNN(NM)                                                                                     \
        if (row_major)                                                                                     \
            hipLaunchKernelGGL((ts_gemm_nn_kernel<V, TB, T, Acc, NM, true>), grid, dim3(kThreads), 0, s, M, N, K, \
                               A, rs, cs, B, ldb, Y, ldy, Acc(alpha), Acc(beta), gate);                          \
        else                                                                                               \
            hipLaunchKernelGGL((ts_gemm_nn_kernel<V, TB, T, Acc, NM, false>), grid, dim3(kThreads), 0, s, M, N, K, \
                               A, rs, cs, B, ldb, Y, ldy, Acc(alpha), Acc(beta), gate);
        if (N <= 1) { AMFLS_TSNN(1) }
        else if (N <= 2) { AMFLS_TSNN(2) }
        else if (N <= 4) { AMFLS_TSNN(4) }
        else if (N <= 8) { AMFLS_TSNN(8) }
        else { AMFLS_TSNN(16) }
#undef AMFLS_TSNN
    } else {
        dim3 grid(static_cast<unsigned>(ceil_div(M, 64)), static_cast<unsigned>(ceil_div(N, 64)));
        if (row_major)
            hipLaunchKernelGGL((gemm_nn_tile_kernel<V, TB, T, Acc, true>), grid, dim3(kThreads), 0, s, M, N,
                               K, A, rs, cs, B, ldb, Y, ldy, Acc(alpha), Acc(beta), gate);
        else
            hipLaunchKernelGGL((gemm_nn_tile_kernel<V, TB, T, Acc, false>), grid, dim3(kThreads), 0, s, M, N,
                               K, A, rs, cs, B, ldb, Y, ldy, Acc(alpha), Acc(beta), gate);
    }
    AMFLS_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ gemm_tn
// Partial sums over a chunk of rows; P[chunk][K][N] (double).
template <typename V, typename TB, bool AROW>
__global__ __launch_bounds__(kThreads) void gemm_tn_tile_kernel(
    int64_t M, int64_t K, int N, const V* __restrict__ A, int64_t rs, int64_t cs,
    const TB* __restrict__ B, int64_t ldb, double* __restrict__ P, int64_t rows_per_chunk, const double* gate) {
    if (gated_off(gate)) return;
    constexpr int BK = 64, BN = 64, BR = 32;
    __shared__ double As[BR][BK + 1];
    __shared__ double Bs[BR][BN];
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * BK;
    const int n0 = blockIdx.y * BN;
    const int64_t chunk = blockIdx.z;
    const int64_t r_begin = chunk * rows_per_chunk;
    const int64_t r_end = std::min<int64_t>(M, r_begin + rows_per_chunk);
    const int t = threadIdx.x, tx = t % 16, ty = t / 16;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int64_t r0 = r_begin; r0 < r_end; r0 += BR) {
#pragma unroll
        for (int l = 0; l < (BR * BK) / kThreads; ++l) {
            const int idx = t + l * kThreads;
            int rr, kk;
            if (AROW) { kk = idx % BK; rr = idx / BK; } else { rr = idx % BR; kk = idx / BR; }
            const int64_t gr = r0 + rr, gk = k0 + kk;
            As[rr][kk] = (gr < r_end && gk < K) ? ld_d(A, gr * rs + gk * cs) : 0.0;
        }
#pragma unroll
        for (int l = 0; l < (BR * BN) / kThreads; ++l) {
            const int idx = t + l * kThreads;
            const int c = idx % BN, rr = idx / BN;
            const int64_t gr = r0 + rr;
            const int gc = n0 + c;
            Bs[rr][c] = (gr < r_end && gc < N) ? static_cast<double>(B[gr * ldb + gc]) : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int rr = 0; rr < BR; ++rr) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[rr][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[rr][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t gk = k0 + ty * 4 + i;
        if (gk >= K) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            P[(chunk * K + gk) * N + gn] = acc[i][j];
        }
    }
}

// Tall-skinny C = A^T B for K * N <= 4096: each block reduces a contiguous
// range of rows into all K x N outputs (per-thread accumulators), staging
// BR x K rows of A and BR x N rows of B through LDS with coalesced loads.
template <typename V, typename TB, int ACC, bool AROW>
__global__ __launch_bounds__(kThreads) void ts_gemm_tn_kernel(
    int64_t M, int K, int N, const V* __restrict__ A, int64_t rs, int64_t cs, const TB* __restrict__ B,
    int64_t ldb, double* __restrict__ P, int64_t rows_per_block, int BR, const double* gate) {
    if (gated_off(gate)) return;
    extern __shared__ double smem[];
    double* As = smem;            // BR x K
    double* Bs = smem + BR * K;   // BR x N
    const int t = threadIdx.x;
    const int KN = K * N;
    const int64_t r_begin = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    const int64_t r_end = std::min<int64_t>(M, r_begin + rows_per_block);
    double acc[ACC];
#pragma unroll
    for (int a = 0; a < ACC; ++a) acc[a] = 0.0;
    for (int64_t r0 = r_begin; r0 < r_end; r0 += BR) {
        const int nr = static_cast<int>(std::min<int64_t>(BR, r_end - r0));
        for (int idx = t; idx < BR * K; idx += kThreads) {
            int rr, kk;
            if (AROW) { rr = idx / K; kk = idx % K; } else { rr = idx % BR; kk = idx / BR; }
            As[rr * K + kk] = rr < nr ? ld_d(A, (r0 + rr) * rs + static_cast<int64_t>(kk) * cs) : 0.0;
        }
        for (int idx = t; idx < BR * N; idx += kThreads) {
            const int rr = idx / N, j = idx % N;
            Bs[rr * N + j] = rr < nr ? static_cast<double>(B[(r0 + rr) * ldb + j]) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < ACC; ++a) {
            const int o = t + a * kThreads;
            if (o < KN) {
                const int kk = o / N, j = o % N;
                double v = acc[a];
                for (int rr = 0; rr < nr; ++rr) v = fma(As[rr * K + kk], Bs[rr * N + j], v);
                acc[a] = v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < ACC; ++a) {
        const int o = t + a * kThreads;
        if (o < KN) P[static_cast<int64_t>(blockIdx.x) * KN + o] = acc[a];
    }
}

// out[o] = alpha * sum_b P[b][o]: one wavefront per output, lanes over blocks.
template <typename TC>
__global__ __launch_bounds__(kThreads) void reduce_partials_kernel(int64_t nb, int64_t KN, int N, const double* P,
                                                                   TC* C, int64_t ldc, double alpha,
                                                                   const double* gate) {
    if (gated_off(gate)) return;
    const int lane = threadIdx.x % kWave;
    const int64_t o = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (o >= KN) return;
    double v = 0.0;
    for (int64_t b = lane; b < nb; b += kWave) v += P[b * KN + o];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) C[(o / N) * ldc + (o % N)] = static_cast<TC>(alpha * v);
}

template <typename V, typename TB>
void gemm_tn_dispatch(hipStream_t s, int64_t M, int64_t K, int N, const V* A, int64_t rs,
                      int64_t cs, const TB* B, int64_t ldb, void* C, ValueType ctype, int64_t ldc,
                      double alpha, Scratch& scratch, const double* gate) {
    if (K <= 0 || N <= 0) return;
    const bool row_major = (cs == 1);
    const int64_t KN = K * N;
    if constexpr (std::is_same_v<V, float> && std::is_same_v<TB, float>) {
        if (row_major && K <= (int64_t{1} << 30) && tall_wide_supported(static_cast<int>(K), N)) {
            tall_tn_wide<float>(s, M, static_cast<int>(K), N, A, rs, B, ldb, C, ctype, ldc, alpha, scratch, gate);
            return;
        }
    }
    if constexpr (std::is_same_v<V, double> && std::is_same_v<TB, double>) {
        if (row_major && K <= kTallMax && N <= kTallMax) {
            int64_t nb = 0, rows_per_block = 0;
            tall_tn_shape(M, nb, rows_per_block);
            scratch.ensure(static_cast<size_t>(nb * KN) * sizeof(double));
            double* P = static_cast<double*>(scratch.ptr);
            tall_tn_f64(s, M, static_cast<int>(K), N, A, rs, B, ldb, P, nb, rows_per_block, gate);
            const dim3 g2(static_cast<unsigned>(ceil_div(KN, kThreads / kWave)));
            if (ctype == ValueType::f64)
                hipLaunchKernelGGL(reduce_partials_kernel<double>, g2, dim3(kThreads), 0, s, nb, KN, N, P,
                                   static_cast<double*>(C), ldc, alpha, gate);
            else
                hipLaunchKernelGGL(reduce_partials_kernel<float>, g2, dim3(kThreads), 0, s, nb, KN, N, P,
                                   static_cast<float*>(C), ldc, alpha, gate);
            AMFLS_LAUNCH_CHECK();
            return;
        }
    }
    if (KN <= 4096 && K <= 4096) {
        // ~512 blocks, each a contiguous row range
        int64_t nb = std::max<int64_t>(1, std::min<int64_t>(512, ceil_div(std::max<int64_t>(M, 1), 32)));
        nb = fit_chunks(scratch, static_cast<size_t>(KN) * sizeof(double), nb);
        const int64_t rows_per_block = ceil_div(std::max<int64_t>(M, 1), nb);
        nb = ceil_div(std::max<int64_t>(M, 1), rows_per_block);
        int BR = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(64, 6144 / (K + N))));
        const size_t lds = static_cast<size_t>(BR) * (K + N) * sizeof(double);
        scratch.ensure(static_cast<size_t>(nb * KN) * sizeof(double));
        double* P = static_cast<double*>(scratch.ptr);
        const int acc = static_cast<int>(ceil_div(KN, kThreads));
        const dim3 grid(static_cast<unsigned>(nb));
#define AMFLS_TSTN(NA)                                                                                       \
        if (row_major)                                                                                       \
            hipLaunchKernelGGL((ts_gemm_tn_kernel<V, TB, NA, true>), grid, dim3(kThreads), lds, s, M,          \
                               static_cast<int>(K), N, A, rs, cs, B, ldb, P, rows_per_block, BR, gate);           \
        else                                                                                                 \
            hipLaunchKernelGGL((ts_gemm_tn_kernel<V, TB, NA, false>), grid, dim3(kThreads), lds, s, M,         \
                               static_cast<int>(K), N, A, rs, cs, B, ldb, P, rows_per_block, BR, gate);
        if (acc <= 1) { AMFLS_TSTN(1) }
        else if (acc <= 2) { AMFLS_TSTN(2) }
        else if (acc <= 4) { AMFLS_TSTN(4) }
        else if (acc <= 8) { AMFLS_TSTN(8) }
        else { AMFLS_TSTN(16) }
#undef AMFLS_TSTN
        AMFLS_LAUNCH_CHECK();
        const dim3 g2(static_cast<unsigned>(ceil_div(KN, kThreads / kWave)));
        if (ctype == ValueType::f64)
            hipLaunchKernelGGL(reduce_partials_kernel<double>, g2, dim3(kThreads), 0, s, nb, KN, N, P,
                               static_cast<double*>(C), ldc, alpha, gate);
        else
            hipLaunchKernelGGL(reduce_partials_kernel<float>, g2, dim3(kThreads), 0, s, nb, KN, N, P,
                               static_cast<float*>(C), ldc, alpha, gate);
        AMFLS_LAUNCH_CHECK();
        return;
    }
    // wide outputs: 64 x 64 output tiles, split over row chunks
    const int64_t col_blocks = ceil_div(K, 64) * ceil_div(N, 64);
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(M, 256), ceil_div(1024, col_blocks)));
    chunks = std::min<int64_t>(chunks, 65535);
    chunks = fit_chunks(scratch, static_cast<size_t>(KN) * sizeof(double), chunks);
    const int64_t rows_per_chunk = std::max<int64_t>(1, ceil_div(std::max<int64_t>(M, 1), chunks));
    chunks = std::max<int64_t>(1, ceil_div(std::max<int64_t>(M, 1), rows_per_chunk));
    scratch.ensure(static_cast<size_t>(chunks * KN) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    dim3 grid(static_cast<unsigned>(ceil_div(K, 64)), static_cast<unsigned>(ceil_div(N, 64)),
              static_cast<unsigned>(chunks));
    if (row_major)
        hipLaunchKernelGGL((gemm_tn_tile_kernel<V, TB, true>), grid, dim3(kThreads), 0, s, M, K, N, A, rs,
                           cs, B, ldb, P, rows_per_chunk, gate);
    else
        hipLaunchKernelGGL((gemm_tn_tile_kernel<V, TB, false>), grid, dim3(kThreads), 0, s, M, K, N, A,
                           rs, cs, B, ldb, P, rows_per_chunk, gate);
    AMFLS_LAUNCH_CHECK();
    const dim3 g2(static_cast<unsigned>(ceil_div(KN, kThreads / kWave)));
    if (ctype == ValueType::f64)
        hipLaunchKernelGGL(reduce_partials_kernel<double>, g2, dim3(kThreads), 0, s, chunks, KN, N, P,
                           static_cast<double*>(C), ldc, alpha, gate);
    else
        hipLaunchKernelGGL(reduce_partials_kernel<float>, g2, dim3(kThreads), 0, s, chunks, KN, N, P,
                           static_cast<float*>(C), ldc, alpha, gate);
    AMFLS_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ rowmul
// Each block owns 32 rows entirely: loads them to LDS, then overwrites.
template <typename T>
__global__ __launch_bounds__(kThreads) void rowmul_kernel(int64_t rows, int w, int r, const T* X,
                                                          int64_t ldx, const double* __restrict__ M,
                                                          T* Y, int64_t ldy) {
    constexpr int BR = 32;
    __shared__ T Xs[BR][128];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * BR;
    for (int idx = threadIdx.x; idx < BR * w; idx += kThreads) {
        const int rr = idx / w, c = idx % w;
        const int64_t gr = r0 + rr;
        Xs[rr][c] = gr < rows ? X[gr * ldx + c] : T(0);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < BR * r; idx += kThreads) {
        const int rr = idx / r, c = idx % r;
        const int64_t gr = r0 + rr;
        if (gr >= rows) continue;
        double acc = 0.0;
        for (int j = 0; j < w; ++j) acc = fma(static_cast<double>(Xs[rr][j]), M[j * r + c], acc);
        Y[gr * ldy + c] = static_cast<T>(acc);
    }
}

// ------------------------------------------------------------------ vectors
template <typename Tin, typename Tout>
__global__ void copy_block_kernel(int64_t rows, int cols, const Tin* X, int64_t ldx, Tout* Y,
                                  int64_t ldy, double scale) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    const int64_t i = idx / cols;
    const int j = static_cast<int>(idx % cols);
    const double v = static_cast<double>(X[i * ldx + j]);
    Y[i * ldy + j] = static_cast<Tout>(scale == 1.0 ? v : scale * v);
}

template <typename T>
__global__ void axpy_block_kernel(int64_t rows, int cols, double a, const T* X, int64_t ldx, T* Y,
                                  int64_t ldy) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    const int64_t i = idx / cols;
    const int j = static_cast<int>(idx % cols);
    T* y = Y + i * ldy + j;
    *y = static_cast<T>(fma(a, static_cast<double>(X[i * ldx + j]), static_cast<double>(*y)));
}

template <typename T>
__global__ void fill_kernel(T* y, int64_t n, T value) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = value;
}

__global__ void sub_kernel(int64_t n, const double* a, const double* b, double* r) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = a[i] - b[i];
}

__global__ void fma_kernel(int64_t n, double alpha, const double* x, const double* z, double* y) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fma(alpha, x[i], z[i]);
}

__global__ void add_kernel(int64_t n, const double* x, double* y, const double* gate) {
    if (gated_off(gate)) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

__global__ void scale_kernel(int64_t n, double alpha, const double* x, double* y) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = alpha * x[i];
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One-launch norm statistics.  Each block reduces its grid-stride share to
// (max |v|, sum (|v| 2^-e_b)^2, nonfinite count) with e_b the frexp exponent
// of the block's own max; the last block to finish (ticket on `counter`)
// rescales every partial sum to the exponent e of the global max, exactly
// since the factors 2^(2(e_b - e)) are powers of two, and writes out[0..2].
// The last block also resets the counter, so consecutive launches on one
// stream reuse it.
//
// One pass: each thread keeps its squares scaled by 2^-e_t, e_t the exponent
// of its own running max, and rescales that sum by a power of two when e_t
// grows; the sums are brought to e_b before they are added.  Power-of-two
// scalings commute with rounding, so the result equals that of squares
// scaled by 2^-e_b from the start (away from subnormal range).
struct ScaledSquares {
    double m = 0.0, acc = 0.0, bad = 0.0;
    int e = -4096;  // below every double exponent
    __device__ __forceinline__ void add(double x) {
        const double v = fabs(x);
        if (!isfinite(v)) {
            bad += 1.0;
            return;
        }
        if (v == 0.0) return;
        m = fmax(m, v);
        int ev = 0;
        (void)frexp(v, &ev);
        const int en = ev > e ? ev : e;
        acc = ldexp(acc, 2 * (e - en));
        e = en;
        const double sv = ldexp(v, -e);
        acc = fma(sv, sv, acc);
    }
};

template <typename T>
__global__ __launch_bounds__(kThreads) void stats_kernel(int64_t rows, int cols, const T* __restrict__ X,
                                                         int64_t ldx, double* partial,
                                                         double* __restrict__ out, unsigned* counter,
                                                         const double* gate) {
    if (gated_off(gate)) return;
    __shared__ double sm[kThreads / kWave][3];
    __shared__ bool last;
    const int64_t total = rows * cols;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    ScaledSquares st;
    if (cols == 1 || ldx == cols) {
        // contiguous: four loads in flight per thread
        const int64_t step = cols == 1 ? ldx : 1;
        int64_t idx = first;
        for (; idx + 3 * stride < total; idx += 4 * stride) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = static_cast<double>(X[(idx + u * stride) * step]);
#pragma unroll
            for (int u = 0; u < 4; ++u) st.add(v[u]);
        }
        for (; idx < total; idx += stride) st.add(static_cast<double>(X[idx * step]));
    } else if (first < total) {
        // strided rows: (row, column) of the element advanced by the stride
        // without a division per element
        const int64_t di = stride / cols;
        const int dj = static_cast<int>(stride - di * cols);
        int64_t i = first / cols;
        int j = static_cast<int>(first - i * cols);
        for (int64_t idx = first; idx < total; idx += stride) {
            st.add(static_cast<double>(X[i * ldx + j]));
            i += di;
            j += dj;
            if (j >= cols) {
                j -= cols;
                ++i;
            }
        }
    }
    const double m = wave_max(st.m);
    const double bad = wave_sum(st.bad);
    if (lane == 0) { sm[wv][0] = m; sm[wv][2] = bad; }
    __syncthreads();
    double bm = 0.0;
#pragma unroll
    for (int k = 0; k < kThreads / kWave; ++k) bm = fmax(bm, sm[k][0]);
    int eb = 0;
    if (bm > 0.0) (void)frexp(bm, &eb);
    double acc = st.acc > 0.0 ? ldexp(st.acc, 2 * (st.e - eb)) : 0.0;
    acc = wave_sum(acc);
    if (lane == 0) sm[wv][1] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ss = 0.0, bb = 0.0;
        for (int k = 0; k < kThreads / kWave; ++k) { ss += sm[k][1]; bb += sm[k][2]; }
        partial[3 * blockIdx.x] = bm;
        partial[3 * blockIdx.x + 1] = ss;
        partial[3 * blockIdx.x + 2] = bb;
        __threadfence();
        const unsigned ticket = atomicAdd(counter, 1u);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int nb = static_cast<int>(gridDim.x);
    double gm = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) gm = fmax(gm, partial[3 * b]);
    gm = wave_max(gm);
    __syncthreads();
    if (lane == 0) sm[wv][0] = gm;
    __syncthreads();
    gm = 0.0;
#pragma unroll
    for (int k = 0; k < kThreads / kWave; ++k) gm = fmax(gm, sm[k][0]);
    int e = 0;
    if (gm > 0.0) (void)frexp(gm, &e);
    double tot = 0.0, tbad = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
        const double pm = partial[3 * b];
        const double ps = partial[3 * b + 1];
        tbad += partial[3 * b + 2];
        if (pm > 0.0) {
            int ebb = 0;
            (void)frexp(pm, &ebb);
            tot += ldexp(ps, 2 * (ebb - e));
        }
    }
    tot = wave_sum(tot);
    tbad = wave_sum(tbad);
    if (lane == 0) { sm[wv][1] = tot; sm[wv][2] = tbad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ss = 0.0, bb = 0.0;
        for (int k = 0; k < kThreads / kWave; ++k) { ss += sm[k][1]; bb += sm[k][2]; }
        out[0] = gm;
        out[1] = ss;
        out[2] = bb;
        *counter = 0u;
    }
}

// gate = 1 when the DGKS criterion asks for a second orthogonalization pass:
// both norms finite and ||after|| < factor * ||before|| (norms from the
// statistics triples exactly as the host forms them).
__device__ double nominal_norm_dev(const double* st) {
    if (st[2] > 0.0) return __builtin_nan("");
    if (st[0] == 0.0) return 0.0;
    int e = 0;
    (void)frexp(st[0], &e);
    return ldexp(sqrt(st[1]), e);
}

__global__ void second_pass_gate_kernel(const double* before, const double* after, double factor, double* gate) {
    if (threadIdx.x != 0) return;
    const double nb = nominal_norm_dev(before), na = nominal_norm_dev(after);
    *gate = (isfinite(nb) && isfinite(na) && na < factor * nb) ? 1.0 : 0.0;
}

// ------------------------------------------------------------------ RNG
__device__ __forceinline__ uint64_t splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

__device__ __forceinline__ double open_uniform(uint64_t bits) {
    return static_cast<double>((bits >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double gaussian_at(uint64_t key, uint64_t index) {
    const double u1 = open_uniform(splitmix64(key ^ (2ULL * index)));
    const double u2 = open_uniform(splitmix64(key ^ (2ULL * index + 1ULL)));
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

// A block owns kGaussRows consecutive rows; its threads take their entries
// in row-major order, so each wavefront's stores cover whole row segments of
// the row-major block (entry (i, j) is still the column-major counter
// index j * rows + i of the generator).
constexpr int kGaussRows = 64;

template <typename T>
__global__ void gaussian_block_kernel(int64_t rows, int cols, uint64_t key, uint64_t first_col, T* Y,
                                      int64_t ldy) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kGaussRows;
    const int nr = static_cast<int>(std::min<int64_t>(kGaussRows, rows - r0));
    for (int e = threadIdx.x; e < nr * cols; e += blockDim.x) {
        const int64_t i = r0 + e / cols;
        const int j = e % cols;
        const uint64_t index = (first_col + static_cast<uint64_t>(j)) * static_cast<uint64_t>(rows) +
                               static_cast<uint64_t>(i);
        Y[i * ldy + j] = static_cast<T>(gaussian_at(key, index));
    }
}

uint64_t host_splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

// ------------------------------------------------------------------ dense stats
template <typename V>
__global__ __launch_bounds__(kThreads) void dense_row_stats_kernel(int64_t M, int64_t N, const V* A,
                                                                   int64_t rs, int64_t cs,
                                                                   double* sum, double* sq,
                                                                   double* misc_partial) {
    // one wavefront per row; lanes along the row
    const int lane = threadIdx.x % kWave;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    double s = 0.0, q = 0.0, mx = 0.0, bad = 0.0;
    if (row < M) {
        for (int64_t j = lane; j < N; j += kWave) {
            const double v = fabs(to_double(A[row * rs + j * cs]));
            if (!isfinite(v)) { bad += 1.0; continue; }
            s += v;
            q = fma(v, v, q);
            mx = fmax(mx, v);
        }
    }
    s = wave_sum(s);
    q = wave_sum(q);
    mx = wave_max(mx);
    bad = wave_sum(bad);
    if (row < M && lane == 0) {
        sum[row] = s;
        sq[row] = q;
        misc_partial[2 * row] = mx;
        misc_partial[2 * row + 1] = bad;
    }
}

template <typename V>
__global__ __launch_bounds__(kThreads) void dense_col_stats_kernel(int64_t M, int64_t N, const V* A,
                                                                   int64_t rs, int64_t cs,
                                                                   double* sum, double* sq) {
    const int lane = threadIdx.x % kWave;
    const int64_t col = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (col >= N) return;
    double s = 0.0, q = 0.0;
    for (int64_t i = lane; i < M; i += kWave) {
        const double v = fabs(to_double(A[i * rs + col * cs]));
        if (!isfinite(v)) continue;
        s += v;
        q = fma(v, v, q);
    }
    s = wave_sum(s);
    q = wave_sum(q);
    if (lane == 0) {
        sum[col] = s;
        sq[col] = q;
    }
}

// Column stats for row-major storage: thread per column, rows looped
// (coalesced along the row).
template <typename V>
__global__ __launch_bounds__(kThreads) void dense_col_stats_rowmajor_kernel(
    int64_t M, int64_t N, const V* A, int64_t rs, double* sum, double* sq, int64_t rows_per_chunk,
    double* psum, double* psq) {
    const int64_t col = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int64_t chunk = blockIdx.y;
    if (col >= N) return;
    const int64_t r0 = chunk * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(M, r0 + rows_per_chunk);
    double s = 0.0, q = 0.0;
    for (int64_t i = r0; i < r1; ++i) {
        const double v = fabs(to_double(A[i * rs + col]));
        if (!isfinite(v)) continue;
        s += v;
        q = fma(v, v, q);
    }
    psum[chunk * N + col] = s;
    psq[chunk * N + col] = q;
}

__global__ void sum_chunks2_kernel(int64_t chunks, int64_t N, const double* psum, const double* psq,
                                   double* sum, double* sq) {
    const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= N) return;
    double s = 0.0, q = 0.0;
    for (int64_t c = 0; c < chunks; ++c) {
        s += psum[c * N + col];
        q += psq[c * N + col];
    }
    sum[col] = s;
    sq[col] = q;
}

__global__ void misc_final_kernel(int64_t rows, const double* misc_partial, double* misc) {
    __shared__ double smax[kThreads / kWave];
    __shared__ double scnt[kThreads / kWave];
    double m = 0.0, bad = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += kThreads) {
        m = fmax(m, misc_partial[2 * r]);
        bad += misc_partial[2 * r + 1];
    }
    m = wave_max(m);
    bad = wave_sum(bad);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) { smax[w] = m; scnt[w] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0, bb = 0.0;
        for (int k = 0; k < kThreads / kWave; ++k) { mm = fmax(mm, smax[k]); bb += scnt[k]; }
        misc[0] = mm;
        misc[1] = bb;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <typename T, typename TB>
void gemm_nn(hipStream_t s, int64_t M, int N, int64_t K, const Strided& A, const TB* B, int64_t ldb,
             T* Y, int64_t ldy, double alpha, double beta, const double* gate) {
    switch (A.type) {
        case ValueType::f64:
            gemm_nn_dispatch<double, TB, T>(s, M, N, K, static_cast<const double*>(A.ptr), A.rs, A.cs, B,
                                            ldb, Y, ldy, alpha, beta, gate);
            break;
        case ValueType::f32:
            gemm_nn_dispatch<float, TB, T>(s, M, N, K, static_cast<const float*>(A.ptr), A.rs, A.cs, B,
                                           ldb, Y, ldy, alpha, beta, gate);
            break;
        case ValueType::bf16:
            gemm_nn_dispatch<__hip_bfloat16, TB, T>(s, M, N, K,
                                                    static_cast<const __hip_bfloat16*>(A.ptr), A.rs,
                                                    A.cs, B, ldb, Y, ldy, alpha, beta, gate);
            break;
    }
}

template <typename TB>
void gemm_tn(hipStream_t s, int64_t M, int64_t K, int N, const Strided& A, const TB* B, int64_t ldb,
             void* C, ValueType ctype, int64_t ldc, double alpha, Scratch& scratch, const double* gate) {
    switch (A.type) {
        case ValueType::f64:
            gemm_tn_dispatch<double, TB>(s, M, K, N, static_cast<const double*>(A.ptr), A.rs, A.cs, B,
                                         ldb, C, ctype, ldc, alpha, scratch, gate);
            break;
        case ValueType::f32:
            gemm_tn_dispatch<float, TB>(s, M, K, N, static_cast<const float*>(A.ptr), A.rs, A.cs, B, ldb,
                                        C, ctype, ldc, alpha, scratch, gate);
            break;
        case ValueType::bf16:
            gemm_tn_dispatch<__hip_bfloat16, TB>(s, M, K, N, static_cast<const __hip_bfloat16*>(A.ptr),
                                                 A.rs, A.cs, B, ldb, C, ctype, ldc, alpha, scratch, gate);
            break;
    }
}

template <typename T>
void rowmul(hipStream_t s, int64_t rows, int w, int r, const T* X, int64_t ldx, const double* M,
            T* Y, int64_t ldy) {
    if (rows <= 0 || r <= 0) return;
    if (w > 128 || r > 128) throw std::invalid_argument("rowmul supports widths up to 128");
    if constexpr (std::is_same_v<T, float>) {
        if (tall_wide_supported(w, r)) {
            tall_nn_wide<float>(s, rows, r, w, X, ldx, M, r, Y, ldy, 1.0, 0.0, nullptr);
            return;
        }
    }
    if constexpr (std::is_same_v<T, double>) {
        // Y = X M with X read whole per tile before Y is written (in place allowed)
        if (w <= kTallMax && r <= kTallMax) {
            tall_nn_f64(s, rows, r, w, X, ldx, M, r, Y, ldy, 1.0, 0.0, nullptr);
            return;
        }
    }
    hipLaunchKernelGGL(rowmul_kernel<T>, dim3(static_cast<unsigned>(ceil_div(rows, 32))), dim3(kThreads), 0,
                       s, rows, w, r, X, ldx, M, Y, ldy);
    AMFLS_LAUNCH_CHECK();
}

template <typename Tin, typename Tout>
void copy_block(hipStream_t s, int64_t rows, int cols, const Tin* X, int64_t ldx, Tout* Y,
                int64_t ldy, double scale) {
    const int64_t total = rows * cols;
    if (total <= 0) return;
    hipLaunchKernelGGL((copy_block_kernel<Tin, Tout>), dim3(static_cast<unsigned>(ceil_div(total, kThreads))),
                       dim3(kThreads), 0, s, rows, cols, X, ldx, Y, ldy, scale);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void fill(hipStream_t s, T* y, int64_t n, T value) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<unsigned>(ceil_div(n, kThreads))), dim3(kThreads), 0, s,
                       y, n, value);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void axpy_block(hipStream_t s, int64_t rows, int cols, double a, const T* X, int64_t ldx, T* Y, int64_t ldy) {
    const int64_t total = rows * cols;
    if (total <= 0) return;
    hipLaunchKernelGGL(axpy_block_kernel<T>, dim3(static_cast<unsigned>(ceil_div(total, kThreads))), dim3(kThreads),
                       0, s, rows, cols, a, X, ldx, Y, ldy);
    AMFLS_LAUNCH_CHECK();
}

void sub_vec(hipStream_t s, int64_t n, const double* a, const double* b, double* r) {
    if (n <= 0) return;
    hipLaunchKernelGGL(sub_kernel, dim3(static_cast<unsigned>(ceil_div(n, kThreads))), dim3(kThreads), 0, s, n,
                       a, b, r);
    AMFLS_LAUNCH_CHECK();
}

void fma_vec(hipStream_t s, int64_t n, double alpha, const double* x, const double* z, double* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fma_kernel, dim3(static_cast<unsigned>(ceil_div(n, kThreads))), dim3(kThreads), 0, s, n,
                       alpha, x, z, y);
    AMFLS_LAUNCH_CHECK();
}

void add_vec(hipStream_t s, int64_t n, const double* x, double* y, const double* gate) {
    if (n <= 0) return;
    hipLaunchKernelGGL(add_kernel, dim3(static_cast<unsigned>(ceil_div(n, kThreads))), dim3(kThreads), 0, s, n,
                       x, y, gate);
    AMFLS_LAUNCH_CHECK();
}

void scale_vec(hipStream_t s, int64_t n, double alpha, const double* x, double* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(scale_kernel, dim3(static_cast<unsigned>(ceil_div(n, kThreads))), dim3(kThreads), 0, s,
                       n, alpha, x, y);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void vector_stats(hipStream_t s, int64_t rows, int cols, const T* X, int64_t ldx, double* out,
                  Scratch& scratch, unsigned* counter, const double* gate) {
    const int64_t total = rows * cols;
    const int blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(1024, ceil_div(total, kThreads * 4))));
    scratch.ensure(static_cast<size_t>(3 * blocks) * sizeof(double));
    hipLaunchKernelGGL(stats_kernel<T>, dim3(blocks), dim3(kThreads), 0, s, rows, cols, X, ldx,
                       static_cast<double*>(scratch.ptr), out, counter, gate);
    AMFLS_LAUNCH_CHECK();
}

void second_pass_gate(hipStream_t s, const double* before, const double* after, double factor, double* gate) {
    hipLaunchKernelGGL(second_pass_gate_kernel, dim3(1), dim3(kWave), 0, s, before, after, factor, gate);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void gaussian_block(hipStream_t s, int64_t rows, int cols, uint64_t seed, uint64_t stream,
                    uint64_t first_col, T* Y, int64_t ldy) {
    const int64_t total = rows * cols;
    if (total <= 0) return;
    const uint64_t key = host_splitmix64(seed) ^ host_splitmix64(stream + 0xd2b74407b1ce6e93ULL);
    hipLaunchKernelGGL(gaussian_block_kernel<T>, dim3(static_cast<unsigned>(ceil_div(rows, kGaussRows))),
                       dim3(kThreads), 0, s, rows, cols, key, first_col, Y, ldy);
    AMFLS_LAUNCH_CHECK();
}

void dense_stats(hipStream_t s, int64_t M, int64_t N, const Strided& A, double* row_sum,
                 double* row_sq, double* col_sum, double* col_sq, double* misc, Scratch& scratch) {
    const bool row_major = A.cs == 1;
    int64_t chunks = 1, rows_per_chunk = M;
    if (row_major) {
        chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(M, 64), ceil_div(4096, ceil_div(N, kThreads))));
        rows_per_chunk = ceil_div(M, chunks);
        chunks = ceil_div(M, rows_per_chunk);
    }
    const size_t misc_bytes = static_cast<size_t>(2 * M) * sizeof(double);
    const size_t chunk_bytes = row_major ? static_cast<size_t>(2 * chunks * N) * sizeof(double) : 0;
    scratch.ensure(misc_bytes + chunk_bytes);
    double* misc_partial = static_cast<double*>(scratch.ptr);
    double* psum = misc_partial + 2 * M;
    double* psq = psum + chunks * N;
    auto run = [&](auto tag) {
        using V = decltype(tag);
        const V* p = static_cast<const V*>(A.ptr);
        hipLaunchKernelGGL(dense_row_stats_kernel<V>, dim3(static_cast<unsigned>(ceil_div(M, kThreads / kWave))),
                           dim3(kThreads), 0, s, M, N, p, A.rs, A.cs, row_sum, row_sq, misc_partial);
        AMFLS_LAUNCH_CHECK();
        if (row_major) {
            hipLaunchKernelGGL(dense_col_stats_rowmajor_kernel<V>,
                               dim3(static_cast<unsigned>(ceil_div(N, kThreads)), static_cast<unsigned>(chunks)),
                               dim3(kThreads), 0, s, M, N, p, A.rs, col_sum, col_sq, rows_per_chunk, psum,
                               psq);
            AMFLS_LAUNCH_CHECK();
            hipLaunchKernelGGL(sum_chunks2_kernel, dim3(static_cast<unsigned>(ceil_div(N, kThreads))),
                               dim3(kThreads), 0, s, chunks, N, psum, psq, col_sum, col_sq);
        } else {
            hipLaunchKernelGGL(dense_col_stats_kernel<V>,
                               dim3(static_cast<unsigned>(ceil_div(N, kThreads / kWave))), dim3(kThreads), 0, s,
                               M, N, p, A.rs, A.cs, col_sum, col_sq);
        }
        AMFLS_LAUNCH_CHECK();
    };
    switch (A.type) {
        case ValueType::f64: run(double{}); break;
        case ValueType::f32: run(float{}); break;
        case ValueType::bf16: run(__hip_bfloat16{}); break;
    }
    hipLaunchKernelGGL(misc_final_kernel, dim3(1), dim3(kThreads), 0, s, M, misc_partial, misc);
    AMFLS_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- instantiations
template void gemm_nn<double, double>(hipStream_t, int64_t, int, int64_t, const Strided&, const double*,
                                      int64_t, double*, int64_t, double, double, const double*);
template void gemm_nn<float, float>(hipStream_t, int64_t, int, int64_t, const Strided&, const float*,
                                    int64_t, float*, int64_t, double, double, const double*);
template void gemm_nn<float, double>(hipStream_t, int64_t, int, int64_t, const Strided&, const double*,
                                     int64_t, float*, int64_t, double, double, const double*);
template void gemm_nn<double, float>(hipStream_t, int64_t, int, int64_t, const Strided&, const float*,
                                     int64_t, double*, int64_t, double, double, const double*);
template void gemm_tn<double>(hipStream_t, int64_t, int64_t, int, const Strided&, const double*, int64_t,
                              void*, ValueType, int64_t, double, Scratch&, const double*);
template void gemm_tn<float>(hipStream_t, int64_t, int64_t, int, const Strided&, const float*, int64_t,
                             void*, ValueType, int64_t, double, Scratch&, const double*);
template void rowmul<double>(hipStream_t, int64_t, int, int, const double*, int64_t, const double*,
                             double*, int64_t);
template void rowmul<float>(hipStream_t, int64_t, int, int, const float*, int64_t, const double*, float*,
                            int64_t);
template void copy_block<double, double>(hipStream_t, int64_t, int, const double*, int64_t, double*,
                                         int64_t, double);
template void copy_block<double, float>(hipStream_t, int64_t, int, const double*, int64_t, float*,
                                        int64_t, double);
template void copy_block<float, double>(hipStream_t, int64_t, int, const float*, int64_t, double*,
                                        int64_t, double);
template void copy_block<float, float>(hipStream_t, int64_t, int, const float*, int64_t, float*, int64_t,
                                       double);
template void axpy_block<double>(hipStream_t, int64_t, int, double, const double*, int64_t, double*, int64_t);
template void axpy_block<float>(hipStream_t, int64_t, int, double, const float*, int64_t, float*, int64_t);
template void fill<double>(hipStream_t, double*, int64_t, double);
template void fill<float>(hipStream_t, float*, int64_t, float);
template void fill<int>(hipStream_t, int*, int64_t, int);
template void vector_stats<double>(hipStream_t, int64_t, int, const double*, int64_t, double*, Scratch&,
                                   unsigned*, const double*);
template void vector_stats<float>(hipStream_t, int64_t, int, const float*, int64_t, double*, Scratch&,
                                  unsigned*, const double*);
template void gaussian_block<double>(hipStream_t, int64_t, int, uint64_t, uint64_t, uint64_t, double*,
                                     int64_t);
template void gaussian_block<float>(hipStream_t, int64_t, int, uint64_t, uint64_t, uint64_t, float*,
                                    int64_t);

}  // namespace amflsx
