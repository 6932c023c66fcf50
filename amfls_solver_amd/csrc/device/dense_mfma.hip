// Dense tall-matrix products on the gfx950 matrix cores.
//
// Y = A X and Z = A^T Y for a row-major A (m x n, stored bf16, fp32 or fp64)
// and row-major blocks X (n x s), Y (m x s) with s <= 64, in three modes:
//
//   split3  A bf16, blocks fp32/fp64: each block value is split into three
//           bf16 terms (hi + mid + lo, 24 significant bits) and multiplied by
//           the exactly representable bf16 A with v_mfma_f32_16x16x32_bf16,
//           fp32 accumulation: fp32-accurate products at bf16 matrix-core
//           rate, bound by the 2 bytes per entry of A.
//   f32     v_mfma_f32_16x16x4_f32 on A and X rounded to fp32.
//   f64     v_mfma_f64_16x16x4_f64 on A upconverted to fp64 (exact for bf16
//           and fp32 storage) and fp64 blocks: fp64-accurate products.
//
// MFMA operands (16 x 16 tiles, wave64): the A operand of lane l holds row
// l % 16, the B operand column l % 16; the reduction index of element t of
// lane group q = l / 16 is assigned the same way in both operands (the K
// dimension is a sum, so any consistent assignment is exact); the 16 x 16
// result of lane l holds column l % 16, rows 4 q + r (r < 4) for the f32 and
// bf16 instructions but rows 4 r + q for v_mfma_f64_16x16x4_f64 (both
// measured on gfx950 with one-hot operands).
//
// Y = A X: one wavefront per 16 rows, K stepped by 32 (eight consecutive
// entries of an A row per lane, one 16-byte load for bf16), X read through
// the cache (it is small); the split-bf16 product on aligned bf16 rows
// stages both the A tile and the X slice of each K step in LDS instead
// (nn_split3_kernel).  Z = A^T Y: a block owns a strip of 64 columns of A
// and a chunk of rows; 32 x 64 tiles of A and 32 x s tiles of Y are staged
// in LDS with coalesced loads; per-chunk partial sums are combined by a
// second kernel in a fixed order (deterministic, no atomics).
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace amflsx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kSplit3 = 0, kF32 = 1, kF64 = 2;
constexpr int kThreads = 256;

template <int MODE> struct AccOf { typedef floatx4 type; };
template <> struct AccOf<kF64> { typedef doublex4 type; };

// Row (within the 16 x 16 tile) of result element r of lane group q.
template <int MODE>
__device__ __forceinline__ int result_row(int q, int r) { return MODE == kF64 ? 4 * r + q : 4 * q + r; }

__device__ __forceinline__ double to_d(double v) { return v; }
__device__ __forceinline__ double to_d(float v) { return static_cast<double>(v); }
__device__ __forceinline__ double to_d(__hip_bfloat16 v) { return static_cast<double>(__bfloat162float(v)); }

__device__ __forceinline__ __bf16 to_bf16(float v) { return static_cast<__bf16>(v); }

// Three-term bf16 split of v: hi = bf16(v), mid = bf16(v - hi), lo = bf16(v - hi - mid).
__device__ __forceinline__ void split3(double v, __bf16& hi, __bf16& mid, __bf16& lo) {
    const float f = static_cast<float>(v);
    hi = to_bf16(f);
    const float r1 = f - static_cast<float>(hi);
    mid = to_bf16(r1);
    lo = to_bf16(r1 - static_cast<float>(mid));
}

// One K=32 step: a[t] (this lane's 8 A-operand values) times b[t] (this
// lane's 8 B-operand values) accumulated into acc.
template <int MODE>
__device__ __forceinline__ void mfma_step(const double* a, const double* b,
                                          typename AccOf<MODE>::type& acc) {
    if constexpr (MODE == kSplit3) {
        bf16x8 av, h, m, l;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            av[t] = static_cast<__bf16>(static_cast<float>(a[t]));  // exact: a is bf16 data
            __bf16 x0, x1, x2;
            split3(b[t], x0, x1, x2);
            h[t] = x0;
            m[t] = x1;
            l[t] = x2;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, l, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, m, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, h, acc, 0, 0, 0);
    } else if constexpr (MODE == kF32) {
#pragma unroll
        for (int t = 0; t < 8; ++t)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(static_cast<float>(a[t]), static_cast<float>(b[t]), acc, 0, 0,
                                                       0);
    } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], acc, 0, 0, 0);
    }
}

// Eight consecutive entries A[row][k0 .. k0+8) (zero past n or past m).
template <typename TA>
__device__ __forceinline__ void load_row8(const TA* __restrict__ A, int64_t lda, int64_t row, int64_t M, int64_t k0,
                                          int64_t N, double* out) {
    if (row < M && k0 + 8 <= N) {
        const TA* p = A + row * lda + k0;
#pragma unroll
        for (int t = 0; t < 8; ++t) out[t] = to_d(p[t]);
    } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) out[t] = (row < M && k0 + t < N) ? to_d(A[row * lda + k0 + t]) : 0.0;
    }
}

// Operand form of the small block X (N x s): for column c and reduction
// index k, element (c, k) at XT[c * nk + k], nk = N rounded up to 64, zero
// padded; split3 stores three bf16 planes (hi, mid, lo), the other modes the
// values in the MFMA input type.
template <typename TX, int MODE>
__global__ void pack_x_kernel(int64_t N, int s, int ncp, int64_t nk, const TX* __restrict__ X, int64_t ldx,
                              void* __restrict__ XT) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(ncp) * nk) return;
    const int c = static_cast<int>(e / nk);
    const int64_t k = e % nk;
    const double v = (c < s && k < N) ? static_cast<double>(X[k * ldx + c]) : 0.0;
    if constexpr (MODE == kSplit3) {
        __bf16* p = static_cast<__bf16*>(XT);
        const int64_t plane = static_cast<int64_t>(ncp) * nk;
        __bf16 h, m, l;
        split3(v, h, m, l);
        p[e] = h;
        p[plane + e] = m;
        p[2 * plane + e] = l;
    } else if constexpr (MODE == kF32) {
        static_cast<float*>(XT)[e] = static_cast<float>(v);
    } else {
        static_cast<double*>(XT)[e] = v;
    }
}

// Eight consecutive stored entries of A (row, k0..k0+8), zero outside.
template <typename TA>
struct Frag8 {
    TA v[8];
    __device__ __forceinline__ void load(const TA* __restrict__ A, int64_t lda, int64_t row, int64_t M, int64_t k0,
                                         int64_t N) {
        if (row < M && k0 + 8 <= N) {
            const TA* p = A + row * lda + k0;
            if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
                // 16-byte vector loads (one for bf16, two for fp32, four for fp64)
                constexpr int W = static_cast<int>(sizeof(TA) * 8 / 16);
                const uint4* src = reinterpret_cast<const uint4*>(p);
                uint4 raw[W];
#pragma unroll
                for (int w = 0; w < W; ++w) raw[w] = src[w];
                __builtin_memcpy(v, raw, sizeof(v));
                return;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) v[t] = p[t];
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) v[t] = (row < M && k0 + t < N) ? A[row * lda + k0 + t] : TA(0.0f);
        }
    }
};

template <typename TA, int MODE>
__device__ __forceinline__ void mfma_packed(const Frag8<TA>& a, const void* __restrict__ XT, int64_t plane,
                                            int64_t off, typename AccOf<MODE>::type& acc) {
    if constexpr (MODE == kSplit3) {
        const bf16x8* x = reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(XT) + off);
        const bf16x8 h = x[0];
        const bf16x8 m = *reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(XT) + plane + off);
        const bf16x8 l = *reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(XT) + 2 * plane + off);
        bf16x8 av;
#pragma unroll
        for (int t = 0; t < 8; ++t) av[t] = static_cast<__bf16>(static_cast<float>(to_d(a.v[t])));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, l, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, m, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, h, acc, 0, 0, 0);
    } else if constexpr (MODE == kF32) {
        const float* x = static_cast<const float*>(XT) + off;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(static_cast<float>(to_d(a.v[t])), x[t], acc, 0, 0, 0);
    } else {
        const double* x = static_cast<const double*>(XT) + off;
#pragma unroll
        for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(to_d(a.v[t]), x[t], acc, 0, 0, 0);
    }
}

// Y (M x s) = A (M x N) X (N x s).  Each wavefront owns RT row tiles of 16
// rows and NT 16-column tiles (s <= 16 NT).  K advances 64 per trip: lane
// (c, q) loads 16 consecutive entries A[row c][k0 + 16 q .. +16) (32 bytes
// of bf16) and feeds them to two K=32 steps, element t of step h standing
// for k = k0 + 16 q + 8 h + t in both operands.  The next trip's A entries
// are loaded before the current trip's matrix-core work.
template <typename TA, typename TX, int MODE, int NT, int RT>
__global__ __launch_bounds__(kThreads) void dense_nn_kernel(int64_t M, int64_t N, int s, const TA* __restrict__ A,
                                                            int64_t lda, const void* __restrict__ XT, int64_t nk,
                                                            int64_t plane, TX* __restrict__ Y, int64_t ldy) {
    using Acc = typename AccOf<MODE>::type;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int c = lane % 16, q = lane / 16;
    const int64_t row0 = (static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + wave) * (16 * RT);
    if (row0 >= M) return;
    Acc acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = Acc{0, 0, 0, 0};
    Frag8<TA> cur[RT][2], nxt[RT][2];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) cur[r][h].load(A, lda, row0 + 16 * r + c, M, 16 * q + 8 * h, N);
    for (int64_t k0 = 0; k0 < N; k0 += 64) {
        if (k0 + 64 < N) {
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int h = 0; h < 2; ++h) nxt[r][h].load(A, lda, row0 + 16 * r + c, M, k0 + 64 + 16 * q + 8 * h, N);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int64_t off = static_cast<int64_t>(16 * j + c) * nk + k0 + 16 * q + 8 * h;
#pragma unroll
                for (int r = 0; r < RT; ++r) mfma_packed<TA, MODE>(cur[r][h], XT, plane, off, acc[r][j]);
            }
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int h = 0; h < 2; ++h) cur[r][h] = nxt[r][h];
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c;
        if (col >= s) continue;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = row0 + 16 * r + result_row<MODE>(q, e);
                if (row < M) Y[row * ldy + col] = static_cast<TX>(acc[r][j][e]);
            }
    }
}

// Direct-to-LDS loads (global_load_lds) and the LDS transpose load of gfx950.
typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(3))) shortx4 lds_shortx4;

// Waits until at most N vector memory operations of this wavefront are outstanding.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Barrier for LDS traffic only (outstanding global loads stay in flight).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Split-bf16 Y = A X for bf16 A (16-byte aligned rows): a block of W
// wavefronts owns 16 RT W rows, a wavefront RT 16-row tiles that share each
// X fragment it reads from LDS.  Each 64-wide K
// trip stages the block's A tile (16 RT W rows x 128 bytes) and the trip's X
// slice (three bf16 planes x 16 NT columns x 64 entries) in LDS with
// coalesced 16-byte loads (eight lanes read the 128 contiguous bytes of a
// row); the global loads of trip t + 3 are in registers while trip t runs
// on the matrix cores, and the two LDS buffers alternate with one barrier
// per trip.  LDS rows are swizzled (slot s of row r holds 16-byte piece
// s ^ (r mod 8)) so the MFMA operand reads of eight consecutive lanes hit
// distinct banks.  Lane (c, q) supplies k = 32 h + 8 q + t of step h for
// both operands.
constexpr int kPipeWaves = 8, kPipeStages = 4;

template <typename TX, int NT, int W = kPipeWaves, int RT = 1>
__global__ __launch_bounds__(64 * W) void nn_split3_kernel(int64_t M, int64_t N, int s,
                                                                    const __hip_bfloat16* __restrict__ A, int64_t lda,
                                                                    const __bf16* __restrict__ XT, int64_t nk,
                                                                    int64_t plane, TX* __restrict__ Y, int64_t ldy) {
    constexpr int NTH = 64 * W, RB = 16 * RT * W, SW = 16 * NT, S = kPipeStages;
    constexpr int APIECES = RB * 8, XPIECES = 3 * SW * 8, PIECES = APIECES + XPIECES;
    constexpr int PPT = (PIECES + NTH - 1) / NTH;
    __shared__ __align__(16) uint4 As[2][APIECES];
    __shared__ __align__(16) uint4 Xs[2][XPIECES];
    const int t = threadIdx.x, lane = t % kWave, wave = t / kWave;
    const int c = lane % 16, q = lane / 16;
    const int64_t rb0 = static_cast<int64_t>(blockIdx.x) * RB;
    const int64_t trips = (N + 63) / 64;
    floatx4 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = floatx4{0, 0, 0, 0};
    uint4 stg[S - 1][PPT];
    auto fetch = [&](uint4 (&dst)[PPT], int64_t k0) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            const int e = t + p * NTH;
            if (e < APIECES) {
                const int row = e / 8, piece = (e % 8) ^ (row & 7);
                const int64_t grow = rb0 + row, kk = k0 + 8 * piece;
                uint4 v = uint4{0, 0, 0, 0};
                if (grow < M && kk + 8 <= N) {
                    v = *reinterpret_cast<const uint4*>(A + grow * lda + kk);
                } else if (grow < M && kk < N) {
                    uint32_t w[4] = {0, 0, 0, 0};
                    for (int u = 0; u < 8; ++u)
                        if (kk + u < N)
                            w[u / 2] |= static_cast<uint32_t>(__builtin_bit_cast(uint16_t, A[grow * lda + kk + u]))
                                        << (16 * (u % 2));
                    v = uint4{w[0], w[1], w[2], w[3]};
                }
                dst[p] = v;
            } else if (PIECES % NTH == 0 || e < PIECES) {
                // X planes are zero padded to nk, a multiple of 64: the slice is always whole
                const int x = e - APIECES, plc = x / 8, col = plc % SW, pl = plc / SW;
                const int piece = (x % 8) ^ (col & 7);
                dst[p] = *reinterpret_cast<const uint4*>(XT + pl * plane + static_cast<int64_t>(col) * nk + k0 +
                                                         8 * piece);
            }
        }
    };
    auto store = [&](const uint4 (&src)[PPT], int buf) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            const int e = t + p * NTH;
            if (e < APIECES) As[buf][e] = src[p];
            else if (PIECES % NTH == 0 || e < PIECES) Xs[buf][e - APIECES] = src[p];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sl = (4 * h + q) ^ (c & 7);
            bf16x8 av[RT];
#pragma unroll
            for (int r = 0; r < RT; ++r)
                av[r] = __builtin_bit_cast(bf16x8, As[buf][(16 * (RT * wave + r) + c) * 8 + sl]);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int col = 16 * j + c;
                const bf16x8 xh = __builtin_bit_cast(bf16x8, Xs[buf][(0 * SW + col) * 8 + sl]);
                const bf16x8 xm = __builtin_bit_cast(bf16x8, Xs[buf][(1 * SW + col) * 8 + sl]);
                const bf16x8 xl = __builtin_bit_cast(bf16x8, Xs[buf][(2 * SW + col) * 8 + sl]);
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r], xl, acc[r][j], 0, 0, 0);
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r], xm, acc[r][j], 0, 0, 0);
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r], xh, acc[r][j], 0, 0, 0);
                }
            }
        }
    };
#pragma unroll
    for (int i = 0; i < S - 1; ++i)
        if (i < trips) fetch(stg[i], 64 * i);
    for (int64_t t0 = 0; t0 < trips; t0 += S - 1) {
#pragma unroll
        for (int u = 0; u < S - 1; ++u) {
            const int64_t tt = t0 + u;
            if (tt < trips) {
                store(stg[u], static_cast<int>(tt & 1));
                __syncthreads();
                if (tt + S - 1 < trips) fetch(stg[u], 64 * (tt + S - 1));
                compute(static_cast<int>(tt & 1));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c;
        if (col >= s) continue;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = rb0 + 16 * (RT * wave + r) + 4 * q + e;
                if (row < M) Y[row * ldy + col] = static_cast<TX>(acc[r][j][e]);
            }
    }
}

// Split-bf16 Y = A X for FP32 blocks of at most 32 columns and N a multiple
// of 8, with every tile moved by direct global-to-LDS loads: a block of four
// wavefronts owns 256 rows, each wavefront four 16-row tiles that share
// every X fragment it reads from LDS (a quarter of the X reads per
// matrix-core operation of nn_split3_kernel with one tile, whose register
// pipeline leaves no room for more tiles).  The A tile (256 rows x 64 K,
// 32 KB) and the X planes' slice (3 planes x 16 NT columns x 64 K) of trip
// t + 2 load into one of three LDS stages while trip t runs, with no
// registers; their 16-byte pieces are swizzled (A: piece ^ row % 8, X:
// piece ^ column % 8) so that the operand reads of 16 lanes spread over the
// banks.  The matrix-core operations and their order are those of
// nn_split3_kernel, so the results are bitwise equal.  Rows past the matrix
// load its last row (their outputs are not written); pieces past column N
// load eight zeros from the X planes' padding (A may hold anything past its
// last column, and NaN times the zero of X is NaN).
constexpr int kNlRows = 256, kNlK = 64, kNlRT = 4;

template <typename TX, int NT>
__global__ __launch_bounds__(kThreads, 1) void nn_lds_kernel(int64_t M, int64_t N, int s,
                                                               const __hip_bfloat16* __restrict__ A, int64_t lda,
                                                               const __bf16* __restrict__ XT, int64_t nk, int64_t plane,
                                                               TX* __restrict__ Y, int64_t ldy) {
    static_assert(NT <= 2 && sizeof(TX) == 4, "FP32 blocks of at most 32 columns");
    constexpr int SW = 16 * NT;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kPieceRow = kNlK / 8;                     // 16-byte pieces per row of a K trip
    constexpr int APIECES = kNlRows * kPieceRow;            // 2048
    constexpr int XPIECES = 3 * SW * kPieceRow;
    constexpr int kALoads = APIECES / kWave / kWaves;       // 1 KB loads per wavefront per trip
    constexpr int kXLoads = (XPIECES / kWave + kWaves - 1) / kWaves;
    constexpr int XSLOTS = kXLoads * kWaves * kWave;        // pieces of an X stage (the tail past XPIECES unused)
    constexpr int kLoads = kALoads + kXLoads;
    static_assert(kALoads * kWaves * kWave == APIECES && XPIECES % kWave == 0, "whole 1 KB loads");
    __shared__ __align__(16) uint4 As0[APIECES];
    __shared__ __align__(16) uint4 As1[APIECES];
    __shared__ __align__(16) uint4 As2[APIECES];
    __shared__ __align__(16) uint4 Xs0[XSLOTS];
    __shared__ __align__(16) uint4 Xs1[XSLOTS];
    __shared__ __align__(16) uint4 Xs2[XSLOTS];
    const int t = threadIdx.x, lane = t % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
    const int c = lane % 16, q = lane / 16;
    const int64_t rb0 = static_cast<int64_t>(blockIdx.x) * kNlRows;
    const int64_t trips = (N + kNlK - 1) / kNlK;
    floatx4 acc[kNlRT][NT];
#pragma unroll
    for (int r = 0; r < kNlRT; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = floatx4{0, 0, 0, 0};
    // wavefront w moves 1 KB pieces w * kALoads + u of the A tile (rows 8 c ..
    // 8 c + 7) and w * kXLoads + u of the X slice
    auto issue = [&](int64_t tr, uint4* as, uint4* xs) {
        const int64_t k0 = tr * kNlK;
#pragma unroll
        for (int u = 0; u < kALoads; ++u) {
            const int ck = wave * kALoads + u;
            const int row = 8 * ck + lane / 8;
            const int piece = (lane % 8) ^ (row & 7);
            const int64_t grow = std::min<int64_t>(rb0 + row, M - 1);
            const int64_t kk = k0 + 8 * piece;
            // past column N: eight zeros of the X planes' padding (N <= nk - 8 then)
            const void* src = kk < N ? static_cast<const void*>(A + grow * lda + kk)
                                     : static_cast<const void*>(XT + (nk - 8));
            __builtin_amdgcn_global_load_lds(src, (lds_void*)(as + ck * kWave), 16, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < kXLoads; ++u) {
            const int ck = wave * kXLoads + u;
            const int x = ck * kWave + lane;
            const int col = (x / kPieceRow) % SW, pl = (x / kPieceRow) / SW;
            const int piece = (x % kPieceRow) ^ (col & 7);
            const __bf16* src = x < XPIECES ? XT + pl * plane + static_cast<int64_t>(col) * nk + k0 + 8 * piece : XT;
            __builtin_amdgcn_global_load_lds(src, (lds_void*)(xs + ck * kWave), 16, 0, 0);
        }
    };
    auto trip = [&](int64_t tr, uint4* as, uint4* xs, uint4* as_next2, uint4* xs_next2) {
        if (tr + 1 < trips)
            wait_vmcnt<kLoads>();  // this wavefront's share of trip tr has landed (trip tr + 1's may not)
        else
            wait_vmcnt<0>();
        lds_barrier();  // every wavefront's share; every wavefront is done with trip tr - 1
        bf16x8 av[2][kNlRT];
        bf16x8 xv[2][NT][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sl = 4 * h + q;
#pragma unroll
            for (int r = 0; r < kNlRT; ++r) {
                const int row = 16 * (kNlRT * wave + r) + c;
                av[h][r] = __builtin_bit_cast(bf16x8, as[row * kPieceRow + (sl ^ (row & 7))]);
            }
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int col = 16 * j + c;
#pragma unroll
                for (int pl = 0; pl < 3; ++pl)
                    xv[h][j][pl] = __builtin_bit_cast(bf16x8, xs[(pl * SW + col) * kPieceRow + (sl ^ (col & 7))]);
            }
        }
        if (tr + 2 < trips) issue(tr + 2, as_next2, xs_next2);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int r = 0; r < kNlRT; ++r) {
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[h][r], xv[h][j][2], acc[r][j], 0, 0, 0);
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[h][r], xv[h][j][1], acc[r][j], 0, 0, 0);
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[h][r], xv[h][j][0], acc[r][j], 0, 0, 0);
                }
    };
    // a first direct-to-LDS load into an array that is never read (see tn_lds_kernel)
    __shared__ __align__(16) uint4 sink[kWave];
    __builtin_amdgcn_global_load_lds(XT, (lds_void*)sink, 16, 0, 0);
    if (trips > 0) issue(0, As0, Xs0);
    if (trips > 1) issue(1, As1, Xs1);
    for (int64_t tr = 0; tr < trips; tr += 3) {
        trip(tr, As0, Xs0, As2, Xs2);
        if (tr + 1 < trips) trip(tr + 1, As1, Xs1, As0, Xs0);
        if (tr + 2 < trips) trip(tr + 2, As2, Xs2, As1, Xs1);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c;
        if (col >= s) continue;
#pragma unroll
        for (int r = 0; r < kNlRT; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = rb0 + 16 * (kNlRT * wave + r) + 4 * q + e;
                if (row < M) Y[row * ldy + col] = static_cast<TX>(acc[r][j][e]);
            }
    }
}

// FP64 blocks, s <= 32: as dense_nn_kernel, but the X^T slice of each
// 64-wide K step (16 NT columns x 64 doubles) is staged in LDS once per
// block and shared by its four wavefronts instead of every wavefront
// reading its B operands through the cache (with fp64 values those reads
// outweighed the A stream).  The next step's slice and A entries are in
// flight while the current step runs on the matrix cores; two LDS buffers
// alternate, one barrier per step.
template <typename TA, typename TX, int NT, int RT>
__global__ __launch_bounds__(kThreads) void dense_nn_f64_kernel(int64_t M, int64_t N, int s, const TA* __restrict__ A,
                                                                int64_t lda, const double* __restrict__ XT,
                                                                int64_t nk, TX* __restrict__ Y, int64_t ldy) {
    constexpr int KP = 64 + 1;                     // padded LDS row (bank spread)
    constexpr int XV = 16 * NT * 64 / kThreads;   // slice values per thread
    __shared__ double Xs[2][16 * NT][KP];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int c = lane % 16, q = lane / 16;
    // no early exit: every wavefront takes part in the barriers (rows past M load zeros)
    const int64_t row0 = (static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + wave) * (16 * RT);
    doublex4 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[r][j] = doublex4{0, 0, 0, 0};
    Frag8<TA> cur[RT][2], nxt[RT][2];
    double xr[XV];
    auto load_x = [&](int64_t k0) {
#pragma unroll
        for (int v = 0; v < XV; ++v) {
            const int e = threadIdx.x + kThreads * v;
            xr[v] = XT[static_cast<int64_t>(e / 64) * nk + k0 + e % 64];
        }
    };
    auto store_x = [&](int b) {
#pragma unroll
        for (int v = 0; v < XV; ++v) {
            const int e = threadIdx.x + kThreads * v;
            Xs[b][e / 64][e % 64] = xr[v];
        }
    };
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) cur[r][h].load(A, lda, row0 + 16 * r + c, M, 16 * q + 8 * h, N);
    load_x(0);
    store_x(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = 0; k0 < N; k0 += 64) {
        const bool more = k0 + 64 < N;
        if (more) {
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int h = 0; h < 2; ++h) nxt[r][h].load(A, lda, row0 + 16 * r + c, M, k0 + 64 + 16 * q + 8 * h, N);
            load_x(k0 + 64);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double* xs = &Xs[buf][16 * j + c][16 * q + 8 * h];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const double xv = xs[t];
#pragma unroll
                    for (int r = 0; r < RT; ++r)
                        acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(to_d(cur[r][h].v[t]), xv, acc[r][j], 0, 0, 0);
                }
            }
        if (more) {
            store_x(buf ^ 1);  // that buffer was last read before the previous barrier
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int h = 0; h < 2; ++h) cur[r][h] = nxt[r][h];
        }
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c;
        if (col >= s) continue;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = row0 + 16 * r + result_row<kF64>(q, e);
                if (row < M) Y[row * ldy + col] = static_cast<TX>(acc[r][j][e]);
            }
    }
}

// Partial Z = A^T Y over one chunk of rows for one 64-column strip of A.
// P[chunk][j][col] (j < N, col < s).  Rows advance RS (32 or 64) per trip:
// the RS x 64 A tile and the RS x s Y tile of the next trip are loaded into
// registers while the current trip's tiles in LDS feed the matrix cores.
template <typename TA, typename TX, int MODE, int NT, int RS>
__global__ __launch_bounds__(kThreads) void dense_tn_kernel(int64_t M, int64_t N, int s, const TA* __restrict__ A,
                                                            int64_t lda, const TX* __restrict__ Yb, int64_t ldy,
                                                            int64_t rows_per_chunk, double* __restrict__ P) {
    using Acc = typename AccOf<MODE>::type;
    constexpr int YW = 16 * NT;
    constexpr int YPT = (RS * YW) / kThreads;  // Y values per thread per trip
    constexpr int APT = RS / 32;               // 8-entry A pieces per thread per trip
    __shared__ TA At[RS][64 + 8];
    __shared__ double Yt[RS][YW + 1];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int c = lane % 16, q = lane / 16;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(M, r0 + rows_per_chunk);
    const int ar = threadIdx.x / 8, ac = (threadIdx.x % 8) * 8;  // this thread's A tile slots
    Acc acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = Acc{0, 0, 0, 0};
    Frag8<TA> an[APT];
    double yn[YPT];
    auto fetch = [&](int64_t i0) {
#pragma unroll
        for (int p = 0; p < APT; ++p) an[p].load(A, lda, i0 + ar + 32 * p, r1, j0 + ac, N);
#pragma unroll
        for (int u = 0; u < YPT; ++u) {
            const int e = threadIdx.x + u * kThreads;
            const int rr = e / YW, cc = e % YW;
            const int64_t row = i0 + rr;
            yn[u] = (row < r1 && cc < s) ? static_cast<double>(Yb[row * ldy + cc]) : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < APT; ++p)
#pragma unroll
            for (int t = 0; t < 8; ++t) At[ar + 32 * p][ac + t] = an[p].v[t];
#pragma unroll
        for (int u = 0; u < YPT; ++u) {
            const int e = threadIdx.x + u * kThreads;
            Yt[e / YW][e % YW] = yn[u];
        }
    };
    if (r0 < r1) {
        fetch(r0);
        store();
    }
    __syncthreads();
    for (int64_t i0 = r0; i0 < r1; i0 += RS) {
        const bool more = i0 + RS < r1;
        if (more) fetch(i0 + RS);
#pragma unroll
        for (int h = 0; h < RS / 32; ++h) {
            double a[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) a[t] = to_d(At[32 * h + 8 * q + t][16 * wave + c]);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                double b[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) b[t] = Yt[32 * h + 8 * q + t][16 * j + c];
                mfma_step<MODE>(a, b, acc[j]);
            }
        }
        __syncthreads();
        if (more) store();
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c;
        if (col >= s) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t jj = j0 + 16 * wave + result_row<MODE>(q, r);
            if (jj < N) P[(static_cast<int64_t>(blockIdx.y) * N + jj) * s + col] = static_cast<double>(acc[j][r]);
        }
    }
}

// Split-bf16 partial Z = A^T Y (bf16 A, FP32 or FP64 blocks Y of at most
// 64 columns) over one chunk of rows for one strip of 256 columns of A:
// P[chunk][j][col].  The reduction runs over rows, so both matrix-core
// operands need eight consecutive rows per lane: each 64-row trip stages
// the A tile transposed in LDS (a thread loads an 8 x 8 sub-tile with
// 16-byte row loads, transposes it in registers, and stores eight 16-byte
// column pieces) and the Y tile as three transposed bf16 planes (split
// once per trip, shared by the four wavefronts).  Wavefront w owns column
// tiles w, w + 4, w + 8, w + 12 of the strip and all NT column tiles of Y;
// the next trip's A and Y are in registers while the current trip runs on
// the matrix cores.  Chunks of one strip set are mapped to one XCD (blocks
// are dispatched to XCDs round robin), so the strips of a chunk read its Y
// rows from the same L2.
constexpr int kTnRows = 64, kTnCols = 256, kTnPad = 8;

__device__ __forceinline__ uint32_t pack_lo(uint32_t a, uint32_t b) { return (a & 0xFFFFu) | (b << 16); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b) { return (a >> 16) | (b & 0xFFFF0000u); }

template <typename TX, int NT>
__global__ __launch_bounds__(kThreads, 2) void tn_split3_kernel(int64_t M, int64_t N, int s,
                                                                  const __hip_bfloat16* __restrict__ A, int64_t lda,
                                                                  const TX* __restrict__ Yb, int64_t ldy,
                                                                  int64_t rows_per_chunk, int strips, int chunks,
                                                                  double* __restrict__ P) {
    constexpr int KP = kTnRows + kTnPad;  // LDS row length (bf16)
    constexpr int SW = 16 * NT;
    constexpr int YPAIRS = (SW * (kTnRows / 8) + kThreads - 1) / kThreads;  // (column, 8-row group) per thread
    __shared__ __align__(16) __bf16 At[kTnCols][KP];
    __shared__ __align__(16) __bf16 Yp[3][SW][KP];
    // block -> (strip, chunk); with chunks a multiple of 8 the strips of one
    // chunk share an XCD
    const int b = blockIdx.x;
    int strip, chunk;
    if (chunks % 8 == 0) {
        const int x = b % 8, q = b / 8;
        strip = q % strips;
        chunk = (q / strips) * 8 + x;
    } else {
        strip = b % strips;
        chunk = b / strips;
    }
    const int64_t j0 = static_cast<int64_t>(strip) * kTnCols;
    const int64_t r0 = static_cast<int64_t>(chunk) * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(M, r0 + rows_per_chunk);
    const int t = threadIdx.x, lane = t % kWave, wave = t / kWave;
    const int c16 = lane % 16, q = lane / 16;
    // this thread's 8 x 8 sub-tile of the A tile: eight consecutive threads
    // take the eight row groups of one column group, so the transposed
    // 16-byte stores of a column are contiguous in LDS (a column-group-major
    // assignment strides them 8 rows apart: 1152 bytes, one bank in two)
    const int bi = t % 8, bj = t / 8;
    const int64_t acol = j0 + 8 * bj;
    floatx4 acc[4][NT];
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[e][j] = floatx4{0, 0, 0, 0};
    uint4 ar[8];
    float yr[YPAIRS][8];
    // Tiles of at most 32 Y columns load without per-row branches: a thread
    // whose column group lies inside the matrix loads its eight rows from
    // in-bounds addresses (rows past the matrix read its last row) and zeroes
    // the rows outside the chunk with selects; a group past the edge holds
    // zeros, and one that straddles it (N not a multiple of 8) gathers
    // element by element.  That branch is per thread and fixed for the
    // kernel.  (Per-row branches cost 8-11% of the A^T Y time at widths up
    // to 32; wider Y tiles keep them, since the selects hold more registers
    // there and took the four-tile kernel to two waves per SIMD.)
    const bool col_full = acol + 8 <= N;
    const bool col_part = !col_full && acol < N;
    auto gather_row = [&](int64_t row) {
        uint32_t w[4] = {0, 0, 0, 0};
        if (row < r1)
            for (int e = 0; e < 8; ++e)
                if (acol + e < N) {
                    const uint16_t bits = __builtin_bit_cast(uint16_t, A[row * lda + acol + e]);
                    w[e / 2] |= static_cast<uint32_t>(bits) << (16 * (e % 2));
                }
        return uint4{w[0], w[1], w[2], w[3]};
    };
    auto fetch = [&](int64_t i0) {
        if constexpr (NT > 2) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t row = i0 + 8 * bi + u;
                ar[u] = (row < r1 && col_full) ? *reinterpret_cast<const uint4*>(A + row * lda + acol)
                                               : gather_row(row);
            }
        } else if (col_full) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t row = i0 + 8 * bi + u;
                const int64_t rl = row < M ? row : M - 1;
                const uint4 v = *reinterpret_cast<const uint4*>(A + rl * lda + acol);
                ar[u] = row < r1 ? v : uint4{0, 0, 0, 0};
            }
        } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) ar[u] = col_part ? gather_row(i0 + 8 * bi + u) : uint4{0, 0, 0, 0};
        }
#pragma unroll
        for (int pp = 0; pp < YPAIRS; ++pp) {
            const int pair = t + pp * kThreads;
            const int col = pair % SW, kg = pair / SW;
            const bool in = kg < kTnRows / 8 && col < s;
            if constexpr (NT > 2) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int64_t row = i0 + 8 * kg + u;
                    yr[pp][u] = (in && row < r1) ? static_cast<float>(Yb[row * ldy + col]) : 0.0f;
                }
            } else {
                const int cl = in ? col : 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int64_t row = i0 + 8 * kg + u;
                    const int64_t rl = row < M ? row : M - 1;
                    const float v = static_cast<float>(Yb[rl * ldy + cl]);
                    yr[pp][u] = (in && row < r1) ? v : 0.0f;
                }
            }
        }
    };
    auto store = [&]() {
        // 8 x 8 transpose: column c of the sub-tile, rows 8 bi .. 8 bi + 8
        const uint32_t* w = reinterpret_cast<const uint32_t*>(ar);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint4 lo, hi;
            lo.x = pack_lo(w[0 * 4 + k], w[1 * 4 + k]);
            lo.y = pack_lo(w[2 * 4 + k], w[3 * 4 + k]);
            lo.z = pack_lo(w[4 * 4 + k], w[5 * 4 + k]);
            lo.w = pack_lo(w[6 * 4 + k], w[7 * 4 + k]);
            hi.x = pack_hi(w[0 * 4 + k], w[1 * 4 + k]);
            hi.y = pack_hi(w[2 * 4 + k], w[3 * 4 + k]);
            hi.z = pack_hi(w[4 * 4 + k], w[5 * 4 + k]);
            hi.w = pack_hi(w[6 * 4 + k], w[7 * 4 + k]);
            *reinterpret_cast<uint4*>(&At[8 * bj + 2 * k][8 * bi]) = lo;
            *reinterpret_cast<uint4*>(&At[8 * bj + 2 * k + 1][8 * bi]) = hi;
        }
#pragma unroll
        for (int pp = 0; pp < YPAIRS; ++pp) {
            const int pair = t + pp * kThreads;
            const int col = pair % SW, kg = pair / SW;
            if (kg >= kTnRows / 8) continue;
            bf16x8 h, m, l;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                __bf16 x0, x1, x2;
                split3(static_cast<double>(yr[pp][u]), x0, x1, x2);
                h[u] = x0;
                m[u] = x1;
                l[u] = x2;
            }
            *reinterpret_cast<bf16x8*>(&Yp[0][col][8 * kg]) = h;
            *reinterpret_cast<bf16x8*>(&Yp[1][col][8 * kg]) = m;
            *reinterpret_cast<bf16x8*>(&Yp[2][col][8 * kg]) = l;
        }
    };
    if (r0 < r1) {
        fetch(r0);
        store();
    }
    __syncthreads();
    for (int64_t i0 = r0; i0 < r1; i0 += kTnRows) {
        const bool more = i0 + kTnRows < r1;
        if (more) fetch(i0 + kTnRows);
#pragma unroll
        for (int h = 0; h < kTnRows / 32; ++h) {
            const int k = 32 * h + 8 * q;
            bf16x8 a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = *reinterpret_cast<const bf16x8*>(&At[16 * (wave + 4 * e) + c16][k]);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const bf16x8 yh = *reinterpret_cast<const bf16x8*>(&Yp[0][16 * j + c16][k]);
                const bf16x8 ym = *reinterpret_cast<const bf16x8*>(&Yp[1][16 * j + c16][k]);
                const bf16x8 yl = *reinterpret_cast<const bf16x8*>(&Yp[2][16 * j + c16][k]);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[e], yl, acc[e][j], 0, 0, 0);
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[e], ym, acc[e][j], 0, 0, 0);
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[e], yh, acc[e][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
        if (more) store();
        __syncthreads();
    }
    // result element r of lane (c16, q): column 16 j + c16 of Y, column
    // 16 (wave + 4 e) + 4 q + r of the strip
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c16;
        if (col >= s) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t jj = j0 + 16 * (wave + 4 * e) + 4 * q + r;
                if (jj < N) P[(static_cast<int64_t>(chunk) * N + jj) * s + col] = static_cast<double>(acc[e][j][r]);
            }
    }
}

// The same partial Z = A^T Y for FP32 Y tiles of at most 32 columns, with
// every tile moved by direct global-to-LDS loads (CDNA4 global_load_lds): the
// same block-to-(strip, chunk) mapping and partials as tn_split3_kernel and
// the same matrix-core operands in the same order (with the same chunks, the
// partials are bitwise equal; the dispatch gives it one round of 256 blocks,
// one per CU, since its 158 KB of LDS hold one block per CU).  The A tile
// of trip t + 3 (16 bytes per lane) and the raw FP32 Y tile of trip t + 2
// (4 bytes per lane) are issued while trip t runs on the matrix cores: four
// A stages, so that up to 96 KB of A per CU are in flight, two raw Y
// stages, no registers and no register transposes.  Each
// A row's 16-byte pieces are XOR-swizzled (piece ^ 2 (row % 4)) so that the
// transpose reads spread over the LDS banks.  The A operand (eight
// consecutive rows of one column per lane) comes from the LDS transpose load
// of gfx950, ds_read_b64_tr_b16: when the 16 lanes of a group read the four
// rows of a 4 x 16 block (lane l: row l / 4, columns 4 (l % 4) .. + 3),
// lane o receives column o of the four rows.  Each trip splits its raw Y
// tile into the three bf16 planes once per block (rows outside the chunk
// and columns past s become zero).  Needs N % 8 == 0 (whole pieces); rows
// past the matrix load its last row and are zeroed in LDS before use, and
// pieces past column N load column 0 of their row (their outputs are not
// written).  Every stage is its own LDS array, so the compiler sees that a
// read of one stage does not wait for the loads into another.
constexpr int kTnStage = kTnRows * kTnCols;  // bf16 per A stage

__device__ __forceinline__ int tn_swz(int row) { return 2 * (row & 3); }

template <typename TX, int NT>
__global__ __launch_bounds__(kThreads, 1) void tn_lds_kernel(int64_t M, int64_t N, int s,
                                                               const __hip_bfloat16* __restrict__ A, int64_t lda,
                                                               const TX* __restrict__ Yb, int64_t ldy,
                                                               int64_t rows_per_chunk, int strips, int chunks,
                                                               double* __restrict__ P) {
    static_assert(NT <= 2 && sizeof(TX) == 4, "FP32 Y tiles of at most 32 columns");
    constexpr int KP = kTnRows + kTnPad;
    constexpr int SW = 16 * NT;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kALoads = kTnStage * 2 / (kWave * 16) / kWaves;  // 16-byte loads per lane per trip (8)
    constexpr int kYLoads = kTnRows * SW / kWave / kWaves;          // 4-byte loads per lane per trip
    constexpr int YPAIRS = (SW * (kTnRows / 8) + kThreads - 1) / kThreads;
    __shared__ __align__(16) __bf16 As0[kTnStage];
    __shared__ __align__(16) __bf16 As1[kTnStage];
    __shared__ __align__(16) __bf16 As2[kTnStage];
    __shared__ __align__(16) __bf16 As3[kTnStage];
    __shared__ __align__(16) float Yr0[kTnRows * SW];
    __shared__ __align__(16) float Yr1[kTnRows * SW];
    __shared__ __align__(16) __bf16 Yp[3][SW][KP];
    const int b = blockIdx.x;
    int strip, chunk;
    if (chunks % 8 == 0) {
        const int x = b % 8, qq = b / 8;
        strip = qq % strips;
        chunk = (qq / strips) * 8 + x;
    } else {
        strip = b % strips;
        chunk = b / strips;
    }
    const int64_t j0 = static_cast<int64_t>(strip) * kTnCols;
    const int64_t r0 = static_cast<int64_t>(chunk) * rows_per_chunk;
    const int64_t r1 = std::min<int64_t>(M, r0 + rows_per_chunk);
    const int64_t trips = r0 < r1 ? (r1 - r0 + kTnRows - 1) / kTnRows : 0;
    const int t = threadIdx.x, lane = t % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
    const int c16 = lane % 16, q = lane / 16;
    floatx4 acc[4][NT];
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[e][j] = floatx4{0, 0, 0, 0};
    // the loads of trip `tr` into one A stage (wavefront w moves 1 KB pieces
    // w * kALoads + u of the A tile, tile rows 2 c, 2 c + 1) and into one raw
    // Y stage (256-byte pieces w * kYLoads + u, 64 / SW rows each)
    auto issue_a = [&](int64_t tr, __bf16* as) {
        const int64_t i0 = r0 + tr * kTnRows;
#pragma unroll
        for (int u = 0; u < kALoads; ++u) {
            const int ck = wave * kALoads + u;
            const int row = 2 * ck + lane / 32;
            const int piece = (lane % 32) ^ tn_swz(row);
            const int64_t grow = std::min<int64_t>(i0 + row, M - 1);
            const int64_t gcol = j0 + 8 * piece;
            __builtin_amdgcn_global_load_lds(A + grow * lda + (gcol < N ? gcol : 0), (lds_void*)(as + ck * 512), 16,
                                             0, 0);
        }
    };
    auto issue_y = [&](int64_t tr, float* yr) {
        const int64_t i0 = r0 + tr * kTnRows;
#pragma unroll
        for (int u = 0; u < kYLoads; ++u) {
            const int ck = wave * kYLoads + u;
            const int e = ck * kWave + lane;  // element of the raw tile, row-major SW wide
            const int row = e / SW, col = e % SW;
            const int64_t grow = std::min<int64_t>(i0 + row, M - 1);
            __builtin_amdgcn_global_load_lds(Yb + grow * ldy + (col < s ? col : 0), (lds_void*)(yr + ck * kWave), 4,
                                             0, 0);
        }
    };
    // split of the raw Y tile into the three bf16 planes (once per block)
    auto split_y = [&](int64_t tr, const float* yr) {
        const int64_t i0 = r0 + tr * kTnRows;
#pragma unroll
        for (int pp = 0; pp < YPAIRS; ++pp) {
            const int pair = t + pp * kThreads;
            const int col = pair % SW, kg = pair / SW;
            if (kg >= kTnRows / 8) continue;
            bf16x8 h, m, l;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool in = col < s && i0 + 8 * kg + u < r1;
                const float v = yr[(8 * kg + u) * SW + col];
                __bf16 x0, x1, x2;
                split3(static_cast<double>(in ? v : 0.0f), x0, x1, x2);
                h[u] = x0;
                m[u] = x1;
                l[u] = x2;
            }
            *reinterpret_cast<bf16x8*>(&Yp[0][col][8 * kg]) = h;
            *reinterpret_cast<bf16x8*>(&Yp[1][col][8 * kg]) = m;
            *reinterpret_cast<bf16x8*>(&Yp[2][col][8 * kg]) = l;
        }
    };
    // rows of the A tile past the matrix (the last trip of the last chunk)
    auto zero_tail = [&](int64_t tr, __bf16* as) {
        const int64_t i0 = r0 + tr * kTnRows;
        if (i0 + kTnRows <= M) return;
        const int first = static_cast<int>(M - i0);
        for (int idx = t; idx < (kTnRows - first) * (kTnCols / 8); idx += kThreads)
            *reinterpret_cast<uint4*>(as + (first + idx / (kTnCols / 8)) * kTnCols + 8 * (idx % (kTnCols / 8))) =
                uint4{0, 0, 0, 0};
    };
    // One trip: its tiles have landed in (as, yr).  Each trip issues, in this
    // order, the Y loads of trip tr + 2 (into yr, once split) and the A loads
    // of trip tr + 3 (into as_next3, the stage trip tr - 1 used), so three
    // A stages are in flight while a trip computes.  Issued after Y(tr), in
    // order: A(tr + 1), then Y(tr + 1) and A(tr + 2) as they exist; those may
    // still be in flight.
    auto trip = [&](int64_t tr, __bf16* as, float* yr, __bf16* as_next3) {
        if (tr + 2 < trips)
            wait_vmcnt<2 * kALoads + kYLoads>();
        else if (tr + 1 < trips)
            wait_vmcnt<kALoads + kYLoads>();
        else
            wait_vmcnt<0>();
        lds_barrier();  // every wavefront's share; every wavefront is done with trip tr - 1
        zero_tail(tr, as);
        split_y(tr, yr);
        lds_barrier();
        bf16x8 a[kTnRows / 32][4];
        bf16x8 yp[kTnRows / 32][NT][3];
#pragma unroll
        for (int h = 0; h < kTnRows / 32; ++h) {
            const int k = 32 * h + 8 * q;
            const int r = k + (c16 >> 2);
            const int off = 4 * (c16 & 1);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int piece = 2 * (wave + 4 * e) + ((c16 & 3) >> 1);
                const __bf16* p1 = as + r * kTnCols + 8 * (piece ^ tn_swz(r)) + off;
                const __bf16* p2 = as + (r + 4) * kTnCols + 8 * (piece ^ tn_swz(r + 4)) + off;
                const shortx4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_shortx4*)(p1));
                const shortx4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_shortx4*)(p2));
                a[h][e] = __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
            }
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int pl = 0; pl < 3; ++pl) yp[h][j][pl] = *reinterpret_cast<const bf16x8*>(&Yp[pl][16 * j + c16][k]);
        }
        if (tr + 2 < trips) issue_y(tr + 2, yr);
        if (tr + 3 < trips) issue_a(tr + 3, as_next3);
#pragma unroll
        for (int h = 0; h < kTnRows / 32; ++h)
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h][e], yp[h][j][2], acc[e][j], 0, 0, 0);
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h][e], yp[h][j][1], acc[e][j], 0, 0, 0);
                    acc[e][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h][e], yp[h][j][0], acc[e][j], 0, 0, 0);
                }
    };
    // The compiler inserts waits for direct-to-LDS loads before LDS accesses
    // it cannot tell apart from their destinations; the loads into the first
    // destination it sees are tracked without that distinction, so every read
    // of that array would wait for every load in flight.  A first load into
    // an array that is never read takes that place, and the waits above
    // (explicit, per trip) are the only ones.
    __shared__ __align__(16) float sink[kWave];
    __builtin_amdgcn_global_load_lds(Yb, (lds_void*)sink, 4, 0, 0);
    // the loads before trip 0, in the order the trips keep: A(0), Y(0), A(1),
    // Y(1), A(2)
    if (trips > 0) issue_a(0, As0);
    if (trips > 0) issue_y(0, Yr0);
    if (trips > 1) issue_a(1, As1);
    if (trips > 1) issue_y(1, Yr1);
    if (trips > 2) issue_a(2, As2);
    // trip tr uses A stage tr % 4 and Y stage tr % 2
    for (int64_t tr = 0; tr < trips; tr += 4) {
        trip(tr, As0, Yr0, As3);
        if (tr + 1 < trips) trip(tr + 1, As1, Yr1, As0);
        if (tr + 2 < trips) trip(tr + 2, As2, Yr0, As1);
        if (tr + 3 < trips) trip(tr + 3, As3, Yr1, As2);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + c16;
        if (col >= s) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int64_t jj = j0 + 16 * (wave + 4 * e) + 4 * q + rr;
                if (jj < N) P[(static_cast<int64_t>(chunk) * N + jj) * s + col] = static_cast<double>(acc[e][j][rr]);
            }
    }
}

// Z[j][col] = sum over chunks of P[chunk][j][col]: a wavefront per output,
// lanes over chunks in order, then the butterfly (a fixed order).
template <typename TX>
__global__ __launch_bounds__(kThreads) void dense_tn_reduce_kernel(int64_t N, int s, int64_t chunks,
                                                                   const double* __restrict__ P,
                                                                   TX* __restrict__ Z, int64_t ldz) {
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

template <typename TA, typename TX, int MODE>
void nn_dispatch(hipStream_t st, int64_t M, int64_t N, int s, const TA* A, int64_t lda, const TX* X, int64_t ldx,
                 TX* Y, int64_t ldy, Scratch& scratch) {
    const int NTt = s <= 16 ? 1 : (s <= 32 ? 2 : (s <= 48 ? 3 : 4));
    const int ncp = 16 * NTt;
    const int64_t nk = ceil_div(N, 64) * 64;
    const int64_t plane = static_cast<int64_t>(ncp) * nk;
    const size_t esize = MODE == kSplit3 ? 3 * sizeof(__bf16) : (MODE == kF32 ? sizeof(float) : sizeof(double));
    scratch.ensure(static_cast<size_t>(plane) * esize);
    void* XT = scratch.ptr;
    hipLaunchKernelGGL((pack_x_kernel<TX, MODE>), dim3(static_cast<unsigned>(ceil_div(plane, kThreads))),
                       dim3(kThreads), 0, st, N, s, ncp, nk, X, ldx, XT);
    AMFLS_LAUNCH_CHECK();
    // row tiles per wavefront: four share each X fragment, two for fp64
    // storage (its A fragments take four times the registers of bf16)
    constexpr int RT = (MODE == kF64 && sizeof(TA) == 8) ? 2 : 4;
    const dim3 grid(static_cast<unsigned>(ceil_div(M, 16 * RT * (kThreads / kWave))));
    if constexpr (MODE == kF64) {
        if (NTt <= 2) {
            if (NTt == 1)
                hipLaunchKernelGGL((dense_nn_f64_kernel<TA, TX, 1, RT>), grid, dim3(kThreads), 0, st, M, N, s, A, lda,
                                   static_cast<const double*>(XT), nk, Y, ldy);
            else
                hipLaunchKernelGGL((dense_nn_f64_kernel<TA, TX, 2, RT>), grid, dim3(kThreads), 0, st, M, N, s, A, lda,
                                   static_cast<const double*>(XT), nk, Y, ldy);
            AMFLS_LAUNCH_CHECK();
            return;
        }
    }
    if constexpr (MODE == kSplit3 && std::is_same_v<TA, __hip_bfloat16>) {
        if constexpr (sizeof(TX) == 4) {
            if ((reinterpret_cast<uintptr_t>(A) & 15) == 0 && lda % 8 == 0 && N % 8 == 0 && NTt <= 2) {
                const dim3 gl(static_cast<unsigned>(ceil_div(M, kNlRows)));
                if (NTt == 1)
                    hipLaunchKernelGGL((nn_lds_kernel<TX, 1>), gl, dim3(kThreads), 0, st, M, N, s, A, lda,
                                       static_cast<const __bf16*>(XT), nk, plane, Y, ldy);
                else
                    hipLaunchKernelGGL((nn_lds_kernel<TX, 2>), gl, dim3(kThreads), 0, st, M, N, s, A, lda,
                                       static_cast<const __bf16*>(XT), nk, plane, Y, ldy);
                AMFLS_LAUNCH_CHECK();
                return;
            }
        }
        if ((reinterpret_cast<uintptr_t>(A) & 15) == 0 && lda % 8 == 0) {
            // two row tiles per wavefront where the block still fits the CU as
            // often (16 and 64 columns: A X 7-9% and 20% faster), one where the
            // larger A tile halves the resident blocks (32 columns: 16% slower)
#define AMFLS_NN3(NT)                                                                                          \
            constexpr int RT = (NT == 1 || NT == 4) ? 2 : 1;                                                    \
            const dim3 g3(static_cast<unsigned>(ceil_div(M, 16 * RT * kPipeWaves)));                           \
            hipLaunchKernelGGL((nn_split3_kernel<TX, NT, kPipeWaves, RT>), g3, dim3(64 * kPipeWaves), 0, st, M, N, s, A, lda, \
                               static_cast<const __bf16*>(XT), nk, plane, Y, ldy)
            if (NTt == 1) { AMFLS_NN3(1); }
            else if (NTt == 2) { AMFLS_NN3(2); }
            else if (NTt == 3) { AMFLS_NN3(3); }
            else { AMFLS_NN3(4); }
#undef AMFLS_NN3
            AMFLS_LAUNCH_CHECK();
            return;
        }
    }
#define AMFLS_NN(NT)                                                                                           \
    hipLaunchKernelGGL((dense_nn_kernel<TA, TX, MODE, NT, RT>), grid, dim3(kThreads), 0, st, M, N, s, A, lda, XT, \
                       nk, plane, Y, ldy)
    if (NTt == 1) { AMFLS_NN(1); }
    else if (NTt == 2) { AMFLS_NN(2); }
    else if (NTt == 3) { AMFLS_NN(3); }
    else { AMFLS_NN(4); }
#undef AMFLS_NN
    AMFLS_LAUNCH_CHECK();
}

template <typename TX>
void tn_split3_dispatch(hipStream_t st, int64_t M, int64_t N, int s, const __hip_bfloat16* A, int64_t lda,
                        const TX* Y, int64_t ldy, TX* Z, int64_t ldz, Scratch& scratch) {
    const int strips = static_cast<int>(ceil_div(N, kTnCols));
    // one round of three resident blocks per CU over 256 CUs (768 blocks:
    // a second, partial round measured 10% slower), whole 64-row trips per
    // chunk, a multiple of 8 chunks for the XCD mapping
    // the direct-to-LDS kernel holds one block per CU: one round of 256 blocks
    const bool use_lds = sizeof(TX) == 4 && N % 8 == 0 && s <= 32;
    int64_t chunks =
        std::max<int64_t>(1, std::min<int64_t>(ceil_div(use_lds ? 256 : 768, strips), ceil_div(M, 4 * kTnRows)));
    chunks = fit_chunks(scratch, static_cast<size_t>(N * s) * sizeof(double), chunks);
    if (chunks >= 8) chunks -= chunks % 8;
    const int64_t rows_per_chunk = ceil_div(ceil_div(M, chunks), kTnRows) * kTnRows;
    chunks = ceil_div(M, rows_per_chunk);
    scratch.ensure(static_cast<size_t>(chunks * N * s) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    const dim3 grid(static_cast<unsigned>(strips * chunks));
    const int ch = static_cast<int>(chunks);
#define AMFLS_TN3(NT)                                                                                             \
    hipLaunchKernelGGL((tn_split3_kernel<TX, NT>), grid, dim3(kThreads), 0, st, M, N, s, A, lda, Y, ldy,          \
                       rows_per_chunk, strips, ch, P)
#define AMFLS_TNL(NT)                                                                                             \
    hipLaunchKernelGGL((tn_lds_kernel<TX, NT>), grid, dim3(kThreads), 0, st, M, N, s, A, lda, Y, ldy,             \
                       rows_per_chunk, strips, ch, P)
    bool lds = false;
    if constexpr (sizeof(TX) == 4) {
        if (use_lds) {
            lds = true;
            if (s <= 16) { AMFLS_TNL(1); }
            else { AMFLS_TNL(2); }
        }
    }
    if (lds) {}
    else if (s <= 16) { AMFLS_TN3(1); }
    else if (s <= 32) { AMFLS_TN3(2); }
    else if (s <= 48) { AMFLS_TN3(3); }
    else { AMFLS_TN3(4); }
#undef AMFLS_TN3
#undef AMFLS_TNL
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(dense_tn_reduce_kernel<TX>, dim3(static_cast<unsigned>(ceil_div(N * s, kThreads / kWave))),
                       dim3(kThreads), 0, st, N, s, chunks, P, Z, ldz);
    AMFLS_LAUNCH_CHECK();
}

template <typename TA, typename TX, int MODE>
void tn_dispatch(hipStream_t st, int64_t M, int64_t N, int s, const TA* A, int64_t lda, const TX* Y, int64_t ldy,
                 TX* Z, int64_t ldz, Scratch& scratch) {
    if constexpr (MODE == kSplit3 && std::is_same_v<TA, __hip_bfloat16>) {
        if ((reinterpret_cast<uintptr_t>(A) & 15) == 0 && lda % 8 == 0) {
            tn_split3_dispatch<TX>(st, M, N, s, A, lda, Y, ldy, Z, ldz, scratch);
            return;
        }
    }
    const int64_t strips = ceil_div(N, 64);
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, strips), ceil_div(M, 256)));
    chunks = fit_chunks(scratch, static_cast<size_t>(N * s) * sizeof(double), chunks);
    const int64_t rows_per_chunk = ceil_div(ceil_div(M, chunks), 64) * 64;
    chunks = ceil_div(M, rows_per_chunk);
    scratch.ensure(static_cast<size_t>(chunks * N * s) * sizeof(double));
    double* P = static_cast<double*>(scratch.ptr);
    const dim3 grid(static_cast<unsigned>(strips), static_cast<unsigned>(chunks));
    // 64-row trips when the LDS tiles fit (A stored in 2 or 4 bytes, s <= 32)
    constexpr int RS = sizeof(TA) <= 4 ? 64 : 32;
#define AMFLS_TN(NT)                                                                                               \
    hipLaunchKernelGGL((dense_tn_kernel<TA, TX, MODE, NT, (NT <= 2 ? RS : 32)>), grid, dim3(kThreads), 0, st, M, N, s, \
                       A, lda, Y, ldy, rows_per_chunk, P)
    if (s <= 16) { AMFLS_TN(1); }
    else if (s <= 32) { AMFLS_TN(2); }
    else if (s <= 48) { AMFLS_TN(3); }
    else { AMFLS_TN(4); }
#undef AMFLS_TN
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(dense_tn_reduce_kernel<TX>, dim3(static_cast<unsigned>(ceil_div(N * s, kThreads / kWave))),
                       dim3(kThreads), 0, st, N, s, chunks, P, Z, ldz);
    AMFLS_LAUNCH_CHECK();
}

template <typename TX, int MODE>
void by_storage_nn(hipStream_t st, const Strided& A, int64_t M, int64_t N, int s, const TX* X, int64_t ldx, TX* Y,
                   int64_t ldy, Scratch& scratch) {
    switch (A.type) {
        case ValueType::bf16:
            nn_dispatch<__hip_bfloat16, TX, MODE>(st, M, N, s, static_cast<const __hip_bfloat16*>(A.ptr), A.rs, X,
                                                  ldx, Y, ldy, scratch);
            break;
        case ValueType::f32:
            if constexpr (MODE != kSplit3)
                nn_dispatch<float, TX, MODE>(st, M, N, s, static_cast<const float*>(A.ptr), A.rs, X, ldx, Y, ldy,
                                             scratch);
            break;
        case ValueType::f64:
            if constexpr (MODE != kSplit3)
                nn_dispatch<double, TX, MODE>(st, M, N, s, static_cast<const double*>(A.ptr), A.rs, X, ldx, Y, ldy,
                                              scratch);
            break;
    }
}

template <typename TX, int MODE>
void by_storage_tn(hipStream_t st, const Strided& A, int64_t M, int64_t N, int s, const TX* Y, int64_t ldy, TX* Z,
                   int64_t ldz, Scratch& scratch) {
    switch (A.type) {
        case ValueType::bf16:
            tn_dispatch<__hip_bfloat16, TX, MODE>(st, M, N, s, static_cast<const __hip_bfloat16*>(A.ptr), A.rs, Y,
                                                  ldy, Z, ldz, scratch);
            break;
        case ValueType::f32:
            if constexpr (MODE != kSplit3)
                tn_dispatch<float, TX, MODE>(st, M, N, s, static_cast<const float*>(A.ptr), A.rs, Y, ldy, Z, ldz,
                                             scratch);
            break;
        case ValueType::f64:
            if constexpr (MODE != kSplit3)
                tn_dispatch<double, TX, MODE>(st, M, N, s, static_cast<const double*>(A.ptr), A.rs, Y, ldy, Z, ldz,
                                              scratch);
            break;
    }
}

}  // namespace

DenseMfmaMode dense_mfma_mode(ValueType storage, bool fp64_blocks, bool reduced) {
    if (fp64_blocks && !reduced) return DenseMfmaMode::f64;
    return storage == ValueType::bf16 ? DenseMfmaMode::split3 : DenseMfmaMode::f32;
}

template <typename TX>
void dense_mfma_nn(hipStream_t s, DenseMfmaMode mode, const Strided& A, int64_t M, int64_t N, int ncols, const TX* X,
                   int64_t ldx, TX* Y, int64_t ldy, Scratch& scratch) {
    if (M <= 0 || ncols <= 0) return;
    if (A.cs != 1) throw std::invalid_argument("matrix-core dense products need row-major storage");
    if (ncols > 64) throw std::invalid_argument("matrix-core dense products support blocks up to 64 columns");
    if (mode == DenseMfmaMode::split3 && A.type != ValueType::bf16)
        throw std::invalid_argument("the split-bf16 product needs bf16 storage");
    switch (mode) {
        case DenseMfmaMode::split3: by_storage_nn<TX, kSplit3>(s, A, M, N, ncols, X, ldx, Y, ldy, scratch); break;
        case DenseMfmaMode::f32: by_storage_nn<TX, kF32>(s, A, M, N, ncols, X, ldx, Y, ldy, scratch); break;
        case DenseMfmaMode::f64: by_storage_nn<TX, kF64>(s, A, M, N, ncols, X, ldx, Y, ldy, scratch); break;
    }
}

template <typename TX>
void dense_mfma_tn(hipStream_t s, DenseMfmaMode mode, const Strided& A, int64_t M, int64_t N, int ncols, const TX* Y,
                   int64_t ldy, TX* Z, int64_t ldz, Scratch& scratch) {
    if (N <= 0 || ncols <= 0) return;
    if (A.cs != 1) throw std::invalid_argument("matrix-core dense products need row-major storage");
    if (ncols > 64) throw std::invalid_argument("matrix-core dense products support blocks up to 64 columns");
    if (mode == DenseMfmaMode::split3 && A.type != ValueType::bf16)
        throw std::invalid_argument("the split-bf16 product needs bf16 storage");
    if (M <= 0) throw std::invalid_argument("matrix-core dense products need a nonempty matrix");
    switch (mode) {
        case DenseMfmaMode::split3: by_storage_tn<TX, kSplit3>(s, A, M, N, ncols, Y, ldy, Z, ldz, scratch); break;
        case DenseMfmaMode::f32: by_storage_tn<TX, kF32>(s, A, M, N, ncols, Y, ldy, Z, ldz, scratch); break;
        case DenseMfmaMode::f64: by_storage_tn<TX, kF64>(s, A, M, N, ncols, Y, ldy, Z, ldz, scratch); break;
    }
}

size_t dense_product_scratch_bound(int64_t M, int64_t N, int max_cols) {
    if (M <= 0 || N <= 0 || max_cols <= 0) return 0;
    const int64_t s = std::min(max_cols, 64);
    const int64_t row_chunks = std::max<int64_t>(1, ceil_div(M, 256));
    // tn_split3_dispatch: at most ceil(768 / strips) chunks of N x s doubles
    const int64_t c3 = std::max<int64_t>(1, std::min<int64_t>(ceil_div(768, ceil_div(N, kTnCols)), row_chunks));
    // tn_dispatch (fp32 / fp64 matrix cores): at most ceil(2048 / strips) chunks
    const int64_t c2 = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, ceil_div(N, 64)), row_chunks));
    // the width-one vector-ALU A^T product: at most ceil(1024 / strips) chunks of N doubles
    const int64_t c1 = std::max<int64_t>(1, std::min<int64_t>(ceil_div(1024, ceil_div(N, 8 * kThreads)), row_chunks));
    // nn_dispatch: the packed X^T planes, 64 columns by N rounded up to 64, at most 8 bytes an entry
    const int64_t planes = 64 * ceil_div(N, 64) * 8;
    const int64_t need = std::max({std::max(c3, c2) * N * s * 8, c1 * N * 8, planes});
    return static_cast<size_t>(need);
}

template void dense_mfma_nn<double>(hipStream_t, DenseMfmaMode, const Strided&, int64_t, int64_t, int, const double*,
                                    int64_t, double*, int64_t, Scratch&);
template void dense_mfma_nn<float>(hipStream_t, DenseMfmaMode, const Strided&, int64_t, int64_t, int, const float*,
                                   int64_t, float*, int64_t, Scratch&);
template void dense_mfma_tn<double>(hipStream_t, DenseMfmaMode, const Strided&, int64_t, int64_t, int, const double*,
                                    int64_t, double*, int64_t, Scratch&);
template void dense_mfma_tn<float>(hipStream_t, DenseMfmaMode, const Strided&, int64_t, int64_t, int, const float*,
                                   int64_t, float*, int64_t, Scratch&);

}  // namespace amflsx
