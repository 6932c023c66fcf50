// Tall-skinny block products on the matrix cores for the local work of wide
// block Golub-Kahan levels: C = A^T B (Gram matrices, CGS2 coefficients) and
// Y = alpha A C + beta Y (basis updates, W R^{-1}) with A, B, Y of M rows
// (M up to 1e6 and more), row-major, FP32 or FP64, and K, N <= 64.
//
// tall_tn: a block owns a range of rows; its four wavefronts take
// interleaved row quads and each accumulates every 16 x 16 output tile.  Lane
// (c, q) loads KT contiguous entries of row 4 u + q (16 lanes read one whole
// row, a 16-byte load each for 64 FP32 columns); products are exact in FP64 and
// accumulated on v_mfma_f64_16x16x4_f64; the wavefront sums are added in
// LDS and the per-block partials in block order by a second kernel
// (deterministic).
//
// tall_nn: a wavefront owns TPW tiles of 16 rows; C (K x N) sits in LDS,
// lane (i, q) holds KL = K/4 contiguous entries of its row (16-byte loads),
// and the K sum runs as KL matrix-core steps in which lane q contributes
// entry KL q + t to both operands (any consistent assignment of the K index
// is exact).  Products run on v_mfma_f64_16x16x4_f64 for FP32 blocks too
// (exact products of FP32 entries and FP64 coefficients, one rounding to
// FP32 at the store): rounding C = R^{-1} or CGS coefficients to FP32 first
// weakens the FP32 factorization.  A wavefront reads all its rows before
// writing any, and distinct wavefronts own distinct rows, so Y may alias A
// (in-place W := W M).
//
// basis_apply / basis_project: y = alpha V t + beta y and c = V^T g for the
// n x k basis V of a search (FP32 or FP64) and FP64 vectors, used by the
// refinement's preconditioner.
#include "kernels.hpp"

#include <algorithm>

namespace amflsx {
namespace {

constexpr int kThreads = 256;
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool gate_off(const double* gate) { return gate != nullptr && *gate == 0.0; }

// KL contiguous entries of row `row` from column k0 (zero outside).
template <typename T, int KL>
__device__ __forceinline__ void load_row_piece(const T* __restrict__ A, int64_t rs, int64_t row, int64_t M, int k0,
                                               int K, T (&v)[KL]) {
    const T* src = A + row * rs + k0;
    if (row < M && k0 + KL <= K && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        constexpr int W = static_cast<int>(KL * sizeof(T) / 16);
        if constexpr (W >= 1) {
            uint4 raw[W];
#pragma unroll
            for (int w = 0; w < W; ++w) raw[w] = reinterpret_cast<const uint4*>(src)[w];
            __builtin_memcpy(v, raw, sizeof(v));
            return;
        }
    }
    if constexpr (KL * sizeof(T) == 8) {
        if (row < M && k0 + KL <= K && (reinterpret_cast<uintptr_t>(src) & 7) == 0) {
            const uint2 raw = *reinterpret_cast<const uint2*>(src);
            __builtin_memcpy(v, &raw, sizeof(v));
            return;
        }
    }
#pragma unroll
    for (int t = 0; t < KL; ++t) v[t] = (row < M && k0 + t < K) ? src[t] : T(0);
}

// f64 16x16x4 result: lane (c, q) holds entry (4 r + q, c) of the tile.
// Lane (c, q) loads the KT contiguous entries KT c .. KT c + KT - 1 of a row
// of A (NT of B) with one vector load where aligned, and supplies entry m to
// tile m, so tile (m, n) holds C[KT i + m][NT j + n] at (i, j).  The four
// wavefronts take interleaved row quads of the block's range and each
// accumulates every tile; the wavefront sums are added in LDS, tile by
// tile, in wavefront order.
template <typename T, int KT, int NT, int U>
__global__ __launch_bounds__(kThreads) void tall_tn_rows_kernel(int64_t M, int K, int N, const T* __restrict__ A,
                                                                int64_t rs, const T* __restrict__ B, int64_t ldb,
                                                                double* __restrict__ P, int64_t rows_per_block,
                                                                const double* gate) {
    if (gate_off(gate)) return;
    constexpr int W = kThreads / kWave;
    __shared__ double red[W][256];
    // slab blockIdx.y of 16 KT columns of A (rows k0 .. of C)
    const int k0 = static_cast<int>(blockIdx.y) * 16 * KT;
    double* out = P + static_cast<int64_t>(blockIdx.x) * K * N + static_cast<int64_t>(k0) * N;
    A += k0;
    K = std::min(K - k0, 16 * KT);
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
        T a[U][KT], b[U][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = r0 + 4 * u + q;
            load_row_piece<T, KT>(A, rs, row, r_end, KT * c, K, a[u]);
            load_row_piece<T, NT>(B, ldb, row, r_end, NT * c, N, b[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int m = 0; m < KT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(static_cast<double>(a[u][m]),
                                                                     static_cast<double>(b[u][n]), acc[m][n], 0, 0, 0);
    }
#pragma unroll
    for (int m = 0; m < KT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][(4 * r + q) * 16 + c] = acc[m][n][r];
            __syncthreads();
            const int o = threadIdx.x;  // entry (o / 16, o % 16) of the tile
            const int kk = KT * (o / 16) + m, j = NT * (o % 16) + n;
            if (kk < K && j < N) {
                double v = red[0][o];
#pragma unroll
                for (int w = 1; w < W; ++w) v += red[w][o];
                out[kk * N + j] = v;
            }
            __syncthreads();
        }
}

template <typename TC>
__global__ __launch_bounds__(kThreads) void tall_reduce_kernel(int64_t nb, int64_t KN, int N, const double* P, TC* C,
                                                               int64_t ldc, double alpha, const double* gate) {
    if (gate_off(gate)) return;
    const int lane = threadIdx.x % kWave;
    const int64_t o = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (o >= KN) return;
    double v = 0.0;
    for (int64_t b = lane; b < nb; b += kWave) v += P[b * KN + o];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) C[(o / N) * ldc + (o % N)] = static_cast<TC>(alpha * v);
}

template <typename T, int KL, int NT, int TPW>
__global__ __launch_bounds__(kThreads) void tall_nn_wide_kernel(int64_t M, int N, int K, const T* A, int64_t rs,
                                                                const double* __restrict__ C, int64_t ldc, T* Y,
                                                                int64_t ldy, double alpha, double beta,
                                                                const double* gate) {
    if (gate_off(gate)) return;
    constexpr int KM = 4 * KL, NC = 16 * NT;
    __shared__ double Cs[KM][NC + 1];
    for (int e = threadIdx.x; e < KM * NC; e += kThreads) {
        const int k = e / NC, n = e % NC;
        Cs[k][n] = (k < K && n < N) ? C[k * ldc + n] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int i = lane % 16, q = lane / 16;
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + wave) * TPW;
    T a[TPW][KL];
#pragma unroll
    for (int p = 0; p < TPW; ++p) load_row_piece<T, KL>(A, rs, (tile0 + p) * 16 + i, M, KL * q, K, a[p]);
#pragma unroll
    for (int p = 0; p < TPW; ++p) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < KL; ++t)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(static_cast<double>(a[p][t]), Cs[KL * q + t][16 * n + i],
                                                           acc, 0, 0, 0);
            const int col = 16 * n + i;
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = (tile0 + p) * 16 + 4 * r + q;
                if (row >= M) continue;
                double v = alpha * acc[r];
                if (beta != 0.0) v = fma(beta, static_cast<double>(Y[row * ldy + col]), v);
                Y[row * ldy + col] = static_cast<T>(v);
            }
        }
    }
}

// As tall_nn_wide_kernel for any K: a wavefront owns one 16-row tile and
// the K sum runs over slabs of 64 (C slab in LDS, one barrier pair per
// slab), for the short products against a whole basis (n rows, k columns).
template <typename T, int NT>
__global__ __launch_bounds__(kThreads) void tall_nn_deep_kernel(int64_t M, int N, int K, const T* A, int64_t rs,
                                                                const double* __restrict__ C, int64_t ldc, T* Y,
                                                                int64_t ldy, double alpha, double beta,
                                                                const double* gate) {
    if (gate_off(gate)) return;
    constexpr int KL = 16, KM = 64, NC = 16 * NT;
    __shared__ double Cs[KM][NC + 1];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int i = lane % 16, q = lane / 16;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + wave;
    f64x4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[n] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KM) {
        __syncthreads();
        for (int e = threadIdx.x; e < KM * NC; e += kThreads) {
            const int k = e / NC, n = e % NC;
            Cs[k][n] = (k0 + k < K && n < N) ? C[static_cast<int64_t>(k0 + k) * ldc + n] : 0.0;
        }
        __syncthreads();
        T a[KL];
        load_row_piece<T, KL>(A, rs, tile * 16 + i, M, k0 + KL * q, K, a);
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int t = 0; t < KL; ++t)
                acc[n] = __builtin_amdgcn_mfma_f64_16x16x4f64(static_cast<double>(a[t]), Cs[KL * q + t][16 * n + i],
                                                              acc[n], 0, 0, 0);
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int col = 16 * n + i;
        if (col >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = tile * 16 + 4 * r + q;
            if (row >= M) continue;
            double v = alpha * acc[n][r];
            if (beta != 0.0) v = fma(beta, static_cast<double>(Y[row * ldy + col]), v);
            Y[row * ldy + col] = static_cast<T>(v);
        }
    }
}

// y = alpha V t + beta y for a row-major V (n x k, FP32 or FP64) and FP64
// vectors: a wavefront per row, lanes striding over the row (coalesced),
// t read through the cache.
template <typename TV>
__global__ __launch_bounds__(kThreads) void basis_apply_kernel(int64_t n, int k, const TV* __restrict__ V, int64_t ldv,
                                                               const double* __restrict__ t, double* __restrict__ y,
                                                               double alpha, double beta) {
    const int lane = threadIdx.x % kWave;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const TV* v = V + row * ldv;
    double acc = 0.0;
    for (int j = lane; j < k; j += kWave) acc = fma(static_cast<double>(v[j]), t[j], acc);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) y[row] = beta == 0.0 ? alpha * acc : fma(beta, y[row], alpha * acc);
}

// Partial c = V^T g over one chunk of rows for one slab of 256 columns:
// thread j of the slab sums V[row][j] g[row] over the chunk's rows in order.
template <typename TV>
__global__ __launch_bounds__(kThreads) void basis_project_kernel(int64_t n, int k, const TV* __restrict__ V,
                                                                 int64_t ldv, const double* __restrict__ g,
                                                                 int64_t rows_per_chunk, double* __restrict__ P) {
    const int j = blockIdx.x * kThreads + threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(n, r0 + rows_per_chunk);
    double acc = 0.0;
    if (j < k) {
        int64_t r = r0;
        for (; r + 4 <= r1; r += 4) {
            double v[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = static_cast<double>(V[(r + u) * ldv + j]);
                w[u] = g[r + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = fma(v[u], w[u], acc);
        }
        for (; r < r1; ++r) acc = fma(static_cast<double>(V[r * ldv + j]), g[r], acc);
        P[static_cast<int64_t>(blockIdx.y) * k + j] = acc;
    }
}

// c[j] = sum over chunks of P[ch][j]: 64 columns per block; the four
// wavefronts take every fourth chunk with eight loads in flight, and their
// partial sums are added in wavefront order (a fixed order).  A thread per
// column summing all chunks one dependent load at a time took 114 us for the
// 512 chunks of a 20000-row basis.
__global__ __launch_bounds__(kThreads) void chunk_sum_kernel(int64_t chunks, int k, const double* __restrict__ P,
                                                             double* __restrict__ c) {
    constexpr int G = kThreads / kWave;
    __shared__ double part[G][kWave];
    const int lane = threadIdx.x % kWave, g = threadIdx.x / kWave;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * kWave + lane;
    double acc = 0.0;
    if (j < k) {
        int64_t ch = g;
        for (; ch + 7 * G < chunks; ch += 8 * G) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = P[(ch + u * G) * k + j];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; ch < chunks; ch += G) acc += P[ch * k + j];
    }
    part[g][lane] = acc;
    __syncthreads();
    if (g == 0 && j < k) {
        double tot = part[0][lane];
#pragma unroll
        for (int w = 1; w < G; ++w) tot += part[w][lane];
        c[j] = tot;
    }
}

}  // namespace

bool tall_wide_supported(int K, int N) { return K >= 1 && N >= 1 && N <= 64; }

template <typename T>
void tall_tn_wide(hipStream_t s, int64_t M, int K, int N, const T* A, int64_t rs, const T* B, int64_t ldb, void* C,
                  ValueType ctype, int64_t ldc, double alpha, Scratch& scratch, const double* gate) {
    if (K <= 0 || N <= 0) return;
    const int kt = static_cast<int>(std::min<int64_t>(4, ceil_div(K, 16))), nt = static_cast<int>(ceil_div(N, 16));
    const int slabs = static_cast<int>(ceil_div(K, 64));  // K > 64: slabs of 64 rows of C on grid.y
    constexpr int U = 4;
    constexpr int64_t trip = 4 * 4 * U;  // four wavefronts of U row quads
    const int64_t KN = static_cast<int64_t>(K) * N;
    int64_t nb = std::max<int64_t>(1, std::min<int64_t>(ceil_div(1024, slabs), ceil_div(std::max<int64_t>(M, 1), 256)));
    nb = fit_chunks(scratch, static_cast<size_t>(KN) * sizeof(double), nb);
    const int64_t rows_per_block = ceil_div(ceil_div(std::max<int64_t>(M, 1), nb), trip) * trip;
    nb = ceil_div(std::max<int64_t>(M, 1), rows_per_block);
    scratch.ensure(static_cast<size_t>(nb * KN) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    const dim3 grid(static_cast<unsigned>(nb), static_cast<unsigned>(slabs));
#define AMFLS_TTN(KT, NT)                                                                                              \
    hipLaunchKernelGGL((tall_tn_rows_kernel<T, KT, NT, U>), grid, dim3(kThreads), 0, s, M, K, N, A, rs, B, ldb, P, \
                       rows_per_block, gate)
    switch (kt * 10 + nt) {
        case 11: AMFLS_TTN(1, 1); break;
        case 12: AMFLS_TTN(1, 2); break;
        case 13: AMFLS_TTN(1, 3); break;
        case 14: AMFLS_TTN(1, 4); break;
        case 21: AMFLS_TTN(2, 1); break;
        case 22: AMFLS_TTN(2, 2); break;
        case 23: AMFLS_TTN(2, 3); break;
        case 24: AMFLS_TTN(2, 4); break;
        case 31: AMFLS_TTN(3, 1); break;
        case 32: AMFLS_TTN(3, 2); break;
        case 33: AMFLS_TTN(3, 3); break;
        case 34: AMFLS_TTN(3, 4); break;
        case 41: AMFLS_TTN(4, 1); break;
        case 42: AMFLS_TTN(4, 2); break;
        case 43: AMFLS_TTN(4, 3); break;
        default: AMFLS_TTN(4, 4); break;
    }
#undef AMFLS_TTN
    AMFLS_LAUNCH_CHECK();
    const dim3 g2(static_cast<unsigned>(ceil_div(KN, kThreads / kWave)));
    if (ctype == ValueType::f64)
        hipLaunchKernelGGL(tall_reduce_kernel<double>, g2, dim3(kThreads), 0, s, nb, KN, N, P, static_cast<double*>(C),
                           ldc, alpha, gate);
    else
        hipLaunchKernelGGL(tall_reduce_kernel<float>, g2, dim3(kThreads), 0, s, nb, KN, N, P, static_cast<float*>(C),
                           ldc, alpha, gate);
    AMFLS_LAUNCH_CHECK();
}

template <typename TV>
void basis_apply(hipStream_t s, int64_t n, int k, const TV* V, int64_t ldv, const double* t, double* y, double alpha,
                 double beta) {
    if (n <= 0) return;
    hipLaunchKernelGGL(basis_apply_kernel<TV>, dim3(static_cast<unsigned>(ceil_div(n, kThreads / kWave))),
                       dim3(kThreads), 0, s, n, k, V, ldv, t, y, alpha, beta);
    AMFLS_LAUNCH_CHECK();
}

template <typename TV>
void basis_project(hipStream_t s, int64_t n, int k, const TV* V, int64_t ldv, const double* g, double* c,
                   Scratch& scratch) {
    if (k <= 0) return;
    const int64_t slabs = ceil_div(k, kThreads);
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 16), ceil_div(512, slabs)));
    chunks = fit_chunks(scratch, static_cast<size_t>(k) * sizeof(double), chunks);
    const int64_t rows_per_chunk = ceil_div(std::max<int64_t>(n, 1), chunks);
    chunks = ceil_div(std::max<int64_t>(n, 1), rows_per_chunk);
    scratch.ensure(static_cast<size_t>(chunks * k) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    hipLaunchKernelGGL(basis_project_kernel<TV>, dim3(static_cast<unsigned>(slabs), static_cast<unsigned>(chunks)),
                       dim3(kThreads), 0, s, n, k, V, ldv, g, rows_per_chunk, P);
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(chunk_sum_kernel, dim3(static_cast<unsigned>(ceil_div(k, kWave))), dim3(kThreads), 0, s, chunks,
                       k, P, c);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void tall_nn_wide(hipStream_t s, int64_t M, int N, int K, const T* A, int64_t rs, const double* C, int64_t ldc, T* Y,
                  int64_t ldy, double alpha, double beta, const double* gate) {
    if (M <= 0 || N <= 0) return;
    const int nt = static_cast<int>(ceil_div(N, 16));
    if (K > 64) {
        const dim3 gd(static_cast<unsigned>(ceil_div(ceil_div(M, 16), kThreads / kWave)));
#define AMFLS_DNN(NT) \
        hipLaunchKernelGGL((tall_nn_deep_kernel<T, NT>), gd, dim3(kThreads), 0, s, M, N, K, A, rs, C, ldc, Y, ldy, alpha, beta, gate)
        if (nt <= 1) { AMFLS_DNN(1); } else if (nt <= 2) { AMFLS_DNN(2); } else if (nt <= 3) { AMFLS_DNN(3); } else { AMFLS_DNN(4); }
#undef AMFLS_DNN
        AMFLS_LAUNCH_CHECK();
        return;
    }
    constexpr int TPW = 2;
    const int64_t waves = ceil_div(ceil_div(M, 16), TPW);
    const dim3 grid(static_cast<unsigned>(ceil_div(waves, kThreads / kWave)));
#define AMFLS_TNN(KL, NT)                                                                                           \
    hipLaunchKernelGGL((tall_nn_wide_kernel<T, KL, NT, TPW>), grid, dim3(kThreads), 0, s, M, N, K, A, rs, C, ldc, Y, \
                       ldy, alpha, beta, gate)
    if (K <= 16) {
        if (nt <= 1) { AMFLS_TNN(4, 1); } else if (nt <= 2) { AMFLS_TNN(4, 2); } else if (nt <= 3) { AMFLS_TNN(4, 3); } else { AMFLS_TNN(4, 4); }
    } else if (K <= 32) {
        if (nt <= 1) { AMFLS_TNN(8, 1); } else if (nt <= 2) { AMFLS_TNN(8, 2); } else if (nt <= 3) { AMFLS_TNN(8, 3); } else { AMFLS_TNN(8, 4); }
    } else {
        if (nt <= 1) { AMFLS_TNN(16, 1); } else if (nt <= 2) { AMFLS_TNN(16, 2); } else if (nt <= 3) { AMFLS_TNN(16, 3); } else { AMFLS_TNN(16, 4); }
    }
#undef AMFLS_TNN
    AMFLS_LAUNCH_CHECK();
}

template void tall_tn_wide<float>(hipStream_t, int64_t, int, int, const float*, int64_t, const float*, int64_t, void*,
                                  ValueType, int64_t, double, Scratch&, const double*);
template void tall_tn_wide<double>(hipStream_t, int64_t, int, int, const double*, int64_t, const double*, int64_t,
                                   void*, ValueType, int64_t, double, Scratch&, const double*);
template void tall_nn_wide<float>(hipStream_t, int64_t, int, int, const float*, int64_t, const double*, int64_t,
                                  float*, int64_t, double, double, const double*);
template void tall_nn_wide<double>(hipStream_t, int64_t, int, int, const double*, int64_t, const double*, int64_t,
                                   double*, int64_t, double, double, const double*);
template void basis_apply<float>(hipStream_t, int64_t, int, const float*, int64_t, const double*, double*, double,
                                 double);
template void basis_apply<double>(hipStream_t, int64_t, int, const double*, int64_t, const double*, double*, double,
                                  double);
template void basis_project<float>(hipStream_t, int64_t, int, const float*, int64_t, const double*, double*, Scratch&);
template void basis_project<double>(hipStream_t, int64_t, int, const double*, int64_t, const double*, double*,
                                    Scratch&);

}  // namespace amflsx
