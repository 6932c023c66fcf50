// Narrow dense products on the vector ALUs: Y = A X and Z = A^T Y for a
// row-major A (bf16 / fp32 / fp64 storage) and blocks of at most 8 columns
// (FP64 or FP32 blocks, accumulated in FP64).
//
// At these widths a product does 2 s flops per entry of A, far below what
// the VALU delivers per byte of HBM, so the kernels only have to stream A
// once at full width; the matrix cores would idle on 16-column tiles that
// are mostly padding.
//
//   narrow_nn: a block owns 64 rows.  Per column chunk (256 bytes of each
//   row) the block loads the 64 x KC tile of A with coalesced 16-byte loads
//   and the KC x S slice of X (the next chunk's loads are in flight while
//   the current one is used), stages both in LDS; thread (row, group) then
//   accumulates its row against a quarter of the chunk, reading X as a
//   wavefront-wide broadcast.  The four group sums of a row are added in
//   order at the end.
//
//   narrow_tn: a block owns 2048 columns (8 per thread) of a chunk of rows.
//   Rows advance 64 at a time: their Y values are staged in LDS and every
//   thread streams its 8 entries of each row (16-byte loads, four rows in
//   flight), accumulating 8 x S sums.  Chunk partials P[chunk][col][j] are
//   added in a fixed order by narrow_reduce_kernel, so results are
//   deterministic.
#include "kernels.hpp"

#include <hip/hip_bf16.h>

namespace amflsx {
namespace {

constexpr int kThreads = 256;

template <typename Acc, typename TA>
__device__ __forceinline__ Acc widen(TA v) {
    if constexpr (sizeof(TA) == 2) return static_cast<Acc>(__bfloat162float(v));
    else return static_cast<Acc>(v);
}

// Eight consecutive entries of row `row` from column k0 (zero outside the
// matrix); 16-byte vector loads when the piece is whole and aligned.
template <typename TA>
__device__ __forceinline__ void load8(const TA* __restrict__ A, int64_t lda, int64_t row, int64_t M, int64_t k0,
                                      int64_t N, TA (&v)[8]) {
    if (row < M && k0 + 8 <= N) {
        const TA* p = A + row * lda + k0;
        if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
            constexpr int W = static_cast<int>(sizeof(TA) * 8 / 16);
            uint4 raw[W];
#pragma unroll
            for (int w = 0; w < W; ++w) raw[w] = reinterpret_cast<const uint4*>(p)[w];
            __builtin_memcpy(v, raw, sizeof(v));
            return;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = p[t];
        return;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = (row < M && k0 + t < N) ? A[row * lda + k0 + t] : TA(0.0f);
}

template <typename TA, typename TX, int S>
__global__ __launch_bounds__(kThreads) void narrow_nn_kernel(int64_t M, int64_t N, int s, const TA* __restrict__ A,
                                                             int64_t lda, const TX* __restrict__ X, int64_t ldx,
                                                             TX* __restrict__ Y, int64_t ldy) {
    using Acc = double;
    constexpr int KC = 256 / static_cast<int>(sizeof(TA));  // columns per chunk (256 bytes of a row)
    constexpr int PIECES = KC / 8;                            // 8-entry pieces per tile row
    constexpr int PT = 64 * PIECES / kThreads;                // pieces per thread per chunk
    constexpr int PAD = sizeof(TA) == 8 ? 1 : (sizeof(TA) == 4 ? 1 : 2);
    __shared__ TA As[64][KC + PAD];
    __shared__ Acc Xs[KC][S];
    __shared__ Acc red[4][64][S];
    const int t = threadIdx.x;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int r = t % 64, g = t / 64;
    Acc p[S];
#pragma unroll
    for (int j = 0; j < S; ++j) p[j] = Acc(0);
    constexpr int XPT = (KC * S + kThreads - 1) / kThreads;  // X values per thread per chunk
    TA nxt[PT][8];
    Acc xn[XPT];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int pc = t + kThreads * i;
            load8<TA>(A, lda, row0 + pc / PIECES, M, k0 + 8 * (pc % PIECES), N, nxt[i]);
        }
#pragma unroll
        for (int i = 0; i < XPT; ++i) {
            const int e = t + kThreads * i;
            const int c = e / S, j = e % S;
            const int64_t k = k0 + c;
            xn[i] = (e < KC * S && k < N && j < s) ? static_cast<Acc>(X[k * ldx + j]) : Acc(0);
        }
    };
    fetch(0);
    for (int64_t k0 = 0; k0 < N; k0 += KC) {
#pragma unroll
        for (int i = 0; i < PT; ++i) {
            const int pc = t + kThreads * i;
#pragma unroll
            for (int u = 0; u < 8; ++u) As[pc / PIECES][8 * (pc % PIECES) + u] = nxt[i][u];
        }
#pragma unroll
        for (int i = 0; i < XPT; ++i) {
            const int e = t + kThreads * i;
            if (e < KC * S) Xs[e / S][e % S] = xn[i];
        }
        __syncthreads();
        if (k0 + KC < N) fetch(k0 + KC);  // in flight during the chunk's arithmetic
#pragma unroll 8
        for (int c = g * (KC / 4); c < (g + 1) * (KC / 4); ++c) {
            const Acc a = widen<Acc>(As[r][c]);
#pragma unroll
            for (int j = 0; j < S; ++j) p[j] = fma(a, Xs[c][j], p[j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < S; ++j) red[g][r][j] = p[j];
    __syncthreads();
    if (t < 64 && row0 + t < M) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (j >= s) break;
            Y[(row0 + t) * ldy + j] = static_cast<TX>(((red[0][t][j] + red[1][t][j]) + red[2][t][j]) + red[3][t][j]);
        }
    }
}

template <typename TA, typename TX, int S>
__global__ __launch_bounds__(kThreads) void narrow_tn_kernel(int64_t M, int64_t N, int s, const TA* __restrict__ A,
                                                             int64_t lda, const TX* __restrict__ Yb, int64_t ldy,
                                                             int64_t rows_per_chunk, double* __restrict__ P) {
    using Acc = double;
    __shared__ Acc Ys[64][S];
    const int t = threadIdx.x;
    const int64_t c0 = static_cast<int64_t>(blockIdx.x) * (8 * kThreads) + 8 * t;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(M, r0 + rows_per_chunk);
    Acc acc[8][S];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int j = 0; j < S; ++j) acc[u][j] = Acc(0);
    for (int64_t i0 = r0; i0 < r1; i0 += 64) {
        const int nr = static_cast<int>(std::min<int64_t>(64, r1 - i0));
        for (int e = t; e < 64 * S; e += kThreads) {
            const int rr = e / S, j = e % S;
            Ys[rr][j] = (rr < nr && j < s) ? static_cast<Acc>(Yb[(i0 + rr) * ldy + j]) : Acc(0);
        }
        __syncthreads();
        if (c0 < N) {
            int rr = 0;
            for (; rr + 3 < nr; rr += 4) {
                TA a[4][8];
#pragma unroll
                for (int q = 0; q < 4; ++q) load8<TA>(A, lda, i0 + rr + q, r1, c0, N, a[q]);
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const Acc av = widen<Acc>(a[q][u]);
#pragma unroll
                        for (int j = 0; j < S; ++j) acc[u][j] = fma(av, Ys[rr + q][j], acc[u][j]);
                    }
            }
            for (; rr < nr; ++rr) {
                TA a[8];
                load8<TA>(A, lda, i0 + rr, r1, c0, N, a);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const Acc av = widen<Acc>(a[u]);
#pragma unroll
                    for (int j = 0; j < S; ++j) acc[u][j] = fma(av, Ys[rr][j], acc[u][j]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int64_t col = c0 + u;
        if (col >= N) break;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (j >= s) break;
            P[(static_cast<int64_t>(blockIdx.y) * N + col) * s + j] = static_cast<double>(acc[u][j]);
        }
    }
}

// Z[j][col] = sum over chunks of P[chunk][j][col]: a wavefront per output,
// lanes over chunks in order, then the butterfly (a fixed order).
template <typename TX>
__global__ __launch_bounds__(kThreads) void narrow_reduce_kernel(int64_t N, int s, int64_t chunks,
                                                                 const double* __restrict__ P, TX* __restrict__ Z,
                                                                 int64_t ldz) {
    const int lane = threadIdx.x % kWave;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (e >= N * s) return;  // whole wavefront
    const int64_t j = e / s;
    const int col = static_cast<int>(e % s);
    double acc = 0.0;
    for (int64_t ch = lane; ch < chunks; ch += kWave) acc += P[(ch * N + j) * s + col];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) Z[j * ldz + col] = static_cast<TX>(acc);
}

template <typename TA, typename TX>
void nn_typed(hipStream_t st, int64_t M, int64_t N, int s, const TA* A, int64_t lda, const TX* X, int64_t ldx, TX* Y,
              int64_t ldy) {
    const dim3 grid(static_cast<unsigned>(ceil_div(M, 64)));
#define AMFLS_NARROW_NN(S)                                                                                        \
    hipLaunchKernelGGL((narrow_nn_kernel<TA, TX, S>), grid, dim3(kThreads), 0, st, M, N, s, A, lda, X, ldx, Y, ldy)
    if (s <= 1) AMFLS_NARROW_NN(1);
    else if (s <= 2) AMFLS_NARROW_NN(2);
    else if (s <= 4) AMFLS_NARROW_NN(4);
    else AMFLS_NARROW_NN(8);
#undef AMFLS_NARROW_NN
    AMFLS_LAUNCH_CHECK();
}

template <typename TA, typename TX>
void tn_typed(hipStream_t st, int64_t M, int64_t N, int s, const TA* A, int64_t lda, const TX* Y, int64_t ldy, TX* Z,
              int64_t ldz, Scratch& scratch) {
    const int64_t strips = ceil_div(N, 8 * kThreads);
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(1024, strips), ceil_div(M, 256)));
    chunks = fit_chunks(scratch, static_cast<size_t>(N * s) * sizeof(double), chunks);
    const int64_t rows_per_chunk = ceil_div(ceil_div(M, chunks), 64) * 64;
    chunks = ceil_div(M, rows_per_chunk);
    scratch.ensure(static_cast<size_t>(chunks * N * s) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    const dim3 grid(static_cast<unsigned>(strips), static_cast<unsigned>(chunks));
#define AMFLS_NARROW_TN(S)                                                                                         \
    hipLaunchKernelGGL((narrow_tn_kernel<TA, TX, S>), grid, dim3(kThreads), 0, st, M, N, s, A, lda, Y, ldy,        \
                       rows_per_chunk, P)
    if (s <= 1) AMFLS_NARROW_TN(1);
    else if (s <= 2) AMFLS_NARROW_TN(2);
    else if (s <= 4) AMFLS_NARROW_TN(4);
    else AMFLS_NARROW_TN(8);
#undef AMFLS_NARROW_TN
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(narrow_reduce_kernel<TX>, dim3(static_cast<unsigned>(ceil_div(N * s, kThreads / kWave))),
                       dim3(kThreads), 0, st, N, s, chunks, P, Z, ldz);
    AMFLS_LAUNCH_CHECK();
}

}  // namespace

bool dense_narrow_supported(const Strided& A, int ncols) {
    return A.cs == 1 && ncols >= 1 && ncols <= 8 &&
           (A.type == ValueType::bf16 || A.type == ValueType::f32 || A.type == ValueType::f64);
}

template <typename TX>
void dense_narrow_nn(hipStream_t st, const Strided& A, int64_t M, int64_t N, int ncols, const TX* X, int64_t ldx,
                     TX* Y, int64_t ldy) {
    if (!dense_narrow_supported(A, ncols)) throw std::invalid_argument("narrow dense product: unsupported layout");
    if (M <= 0 || N <= 0) return;
    switch (A.type) {
        case ValueType::bf16:
            nn_typed<__hip_bfloat16, TX>(st, M, N, ncols, static_cast<const __hip_bfloat16*>(A.ptr), A.rs, X, ldx, Y,
                                         ldy);
            break;
        case ValueType::f32:
            nn_typed<float, TX>(st, M, N, ncols, static_cast<const float*>(A.ptr), A.rs, X, ldx, Y, ldy);
            break;
        default:
            nn_typed<double, TX>(st, M, N, ncols, static_cast<const double*>(A.ptr), A.rs, X, ldx, Y, ldy);
            break;
    }
}

template <typename TX>
void dense_narrow_tn(hipStream_t st, const Strided& A, int64_t M, int64_t N, int ncols, const TX* Y, int64_t ldy,
                     TX* Z, int64_t ldz, Scratch& scratch) {
    if (!dense_narrow_supported(A, ncols)) throw std::invalid_argument("narrow dense product: unsupported layout");
    if (M <= 0 || N <= 0) return;
    switch (A.type) {
        case ValueType::bf16:
            tn_typed<__hip_bfloat16, TX>(st, M, N, ncols, static_cast<const __hip_bfloat16*>(A.ptr), A.rs, Y, ldy, Z,
                                         ldz, scratch);
            break;
        case ValueType::f32:
            tn_typed<float, TX>(st, M, N, ncols, static_cast<const float*>(A.ptr), A.rs, Y, ldy, Z, ldz, scratch);
            break;
        default:
            tn_typed<double, TX>(st, M, N, ncols, static_cast<const double*>(A.ptr), A.rs, Y, ldy, Z, ldz, scratch);
            break;
    }
}

template void dense_narrow_nn<double>(hipStream_t, const Strided&, int64_t, int64_t, int, const double*, int64_t,
                                      double*, int64_t);
template void dense_narrow_nn<float>(hipStream_t, const Strided&, int64_t, int64_t, int, const float*, int64_t,
                                     float*, int64_t);
template void dense_narrow_tn<double>(hipStream_t, const Strided&, int64_t, int64_t, int, const double*, int64_t,
                                      double*, int64_t, Scratch&);
template void dense_narrow_tn<float>(hipStream_t, const Strided&, int64_t, int64_t, int, const float*, int64_t,
                                     float*, int64_t, Scratch&);

}  // namespace amflsx
