// Rank-one Golub-Kahan step kernels (gfx950).
//
// While the block Golub-Kahan frontier is one column wide, a level is a
// sequence of vector operations on tall vectors: project a residual onto a
// basis (c = B^T w), update it (w -= B c), measure it, decide its rank and
// normalize it.  These kernels fuse each operation with its reduction (per-
// block partials plus a last-block combine, so one launch each), take their
// scalars from device memory and honour device gates, so that whole levels
// can be enqueued without the host reading anything back in between.  The
// decisions they take are the host's rank rule (reference
// krylov_workspace.cpp:82-310) evaluated on device-resident statistics.
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace amflsx {
namespace {

using dev::kBlock;

__device__ __forceinline__ bool gates_off(const double* a, const double* b) {
    return dev::gated_off(a) || dev::gated_off(b);
}

// c[o] = sum_i B[i, o] w[i] (o < k <= kBlock) over rows [0, N).  Block b owns
// rows [b * rpb, (b + 1) * rpb) and writes its k sums to partial[b * k + o];
// the last block adds the partials in block order and writes c and coeff
// (coeff = c, or coeff += c when accumulate is set).  Wider bases go through
// project_wide_kernel and project_combine_kernel.
template <typename TB, typename TW>
__global__ __launch_bounds__(kBlock) void project_kernel(int64_t N, int k, const TB* __restrict__ B, int64_t ldb,
                                                         const TW* __restrict__ w, int64_t ldw, int64_t rpb,
                                                         double* partial, double* __restrict__ c,
                                                         double* __restrict__ coeff, int accumulate,
                                                         unsigned* counter, const double* gate_a,
                                                         const double* gate_b) {
    if (gates_off(gate_a, gate_b)) return;
    __shared__ double red[kBlock];
    __shared__ bool last;
    const int t = threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
    const int64_t r1 = std::min<int64_t>(N, r0 + rpb);
    if (k <= kBlock) {
        int kp = 1;
        while (kp < k) kp <<= 1;
        const int S = kBlock / kp;
        const int o = t % kp, sidx = t / kp;
        double acc = 0.0;
        if (o < k) {
            // eight rows per trip: the loads are issued together, the sum
            // keeps row order
            int64_t i = r0 + sidx;
            for (; i + 7 * S < r1; i += 8 * S) {
                double bv[8], wv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    bv[u] = static_cast<double>(B[(i + u * S) * ldb + o]);
                    wv[u] = static_cast<double>(w[(i + u * S) * ldw]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = fma(bv[u], wv[u], acc);
            }
            for (; i < r1; i += S) acc = fma(static_cast<double>(B[i * ldb + o]), static_cast<double>(w[i * ldw]), acc);
        }
        const double v = dev::reduce_shared_outputs(acc, kp, red);
        if (t < k) partial[static_cast<int64_t>(blockIdx.x) * k + t] = v;
    }
    __threadfence();
    __syncthreads();
    if (t == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int nb = static_cast<int>(gridDim.x);
    if (k <= kBlock) {
        int kp = 1;
        while (kp < k) kp <<= 1;
        const int S = kBlock / kp;
        const int o = t % kp, sidx = t / kp;
        double acc = 0.0;
        if (o < k) {
            // plain loads: the acquire fence above made the partials visible
            int b = sidx;
            for (; b + 3 * S < nb; b += 4 * S) {
                double pv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) pv[u] = partial[static_cast<int64_t>(b + u * S) * k + o];
#pragma unroll
                for (int u = 0; u < 4; ++u) acc += pv[u];
            }
            for (; b < nb; b += S) acc += partial[static_cast<int64_t>(b) * k + o];
        }
        const double v = dev::reduce_shared_outputs(acc, kp, red);
        if (t < k) {
            c[t] = v;
            coeff[t] = accumulate ? coeff[t] + v : v;
        }
    }
    if (t == 0) *counter = 0u;
}

// Wide bases (k > kBlock): partial[y][o] = sum over the rows of row chunk y of
// B[i, o] w[i]; block (x, y) owns columns [256 x, 256 x + 256) of chunk y, so
// the grid has many blocks even when N is small; eight rows' loads are in
// flight per trip and the sum keeps row order.
template <typename TB, typename TW>
__global__ __launch_bounds__(kBlock) void project_wide_kernel(int64_t N, int k, const TB* __restrict__ B, int64_t ldb,
                                                              const TW* __restrict__ w, int64_t ldw, int64_t rpb,
                                                              double* __restrict__ partial, const double* gate_a,
                                                              const double* gate_b) {
    if (gates_off(gate_a, gate_b)) return;
    const int o = blockIdx.x * kBlock + threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * rpb;
    const int64_t r1 = std::min<int64_t>(N, r0 + rpb);
    if (o >= k) return;
    double acc = 0.0;
    int64_t i = r0;
    for (; i + 8 <= r1; i += 8) {
        double bv[8], wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            bv[u] = static_cast<double>(B[(i + u) * ldb + o]);
            wv[u] = static_cast<double>(w[(i + u) * ldw]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fma(bv[u], wv[u], acc);
    }
    for (; i < r1; ++i) acc = fma(static_cast<double>(B[i * ldb + o]), static_cast<double>(w[i * ldw]), acc);
    partial[static_cast<int64_t>(blockIdx.y) * k + o] = acc;
}

// c[o] = sum over blocks of partial[b][o] for k > kBlock: a block per 32
// outputs; thread (o, g) adds blocks g, g + 8, ... in order, then the eight
// group sums are added in group order.
constexpr int kCombineCols = 32, kCombineGroups = kBlock / kCombineCols;
__global__ __launch_bounds__(kBlock) void project_combine_kernel(int nb, int k, const double* __restrict__ partial,
                                                                 double* __restrict__ c, double* __restrict__ coeff,
                                                                 int accumulate, const double* gate_a,
                                                                 const double* gate_b) {
    if (gates_off(gate_a, gate_b)) return;
    __shared__ double red[kCombineGroups][kCombineCols];
    const int ol = threadIdx.x % kCombineCols, g = threadIdx.x / kCombineCols;
    const int o = blockIdx.x * kCombineCols + ol;
    double acc = 0.0;
    if (o < k) {
        int b = g;
        for (; b + 3 * kCombineGroups < nb; b += 4 * kCombineGroups) {
            double pv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pv[u] = partial[static_cast<int64_t>(b + u * kCombineGroups) * k + o];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += pv[u];
        }
        for (; b < nb; b += kCombineGroups) acc += partial[static_cast<int64_t>(b) * k + o];
    }
    red[g][ol] = acc;
    __syncthreads();
    if (g == 0 && o < k) {
        double v = red[0][ol];
#pragma unroll
        for (int q = 1; q < kCombineGroups; ++q) v += red[q][ol];
        c[o] = v;
        coeff[o] = accumulate ? coeff[o] + v : v;
    }
}

// w[i] = w[i] - sum_o B[i, o] c[o] (rounded once to T from double), then the
// statistics triple of the updated w into out[0..2].  When gate_out is given
// the last block also sets *gate_out = 1 if ||w_new|| < factor * ||before||
// with both norms finite (the DGKS test for a second pass), else 0.
// NARROW: one thread per row (k <= 16); otherwise one wavefront per row with
// lanes over the k coefficients.
template <typename T, bool NARROW>
__global__ __launch_bounds__(kBlock) void update_kernel(int64_t N, int k, const T* __restrict__ B, int64_t ldb,
                                                        const double* __restrict__ c, T* w, int64_t ldw, int rpb,
                                                        double* partial, double* out, unsigned* counter,
                                                        const double* gate_a, const double* gate_b,
                                                        const double* before, double factor, double* gate_out) {
    if (gates_off(gate_a, gate_b)) return;
    extern __shared__ double vals[];  // rpb updated values
    __shared__ double sm[dev::kWaves];
    __shared__ double cs[16];
    const int t = threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
    const int nr = static_cast<int>(std::min<int64_t>(rpb, N - r0));
    if constexpr (NARROW) {
        if (t < k) cs[t] = c[t];
        __syncthreads();
        int rr = t;
        for (; rr + 3 * kBlock < nr; rr += 4 * kBlock) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0}, wo[4];
            for (int o = 0; o < k; ++o) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[u] = fma(static_cast<double>(B[(r0 + rr + u * kBlock) * ldb + o]), cs[o], acc[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) wo[u] = static_cast<double>(w[(r0 + rr + u * kBlock) * ldw]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T nv = static_cast<T>(wo[u] - acc[u]);
                w[(r0 + rr + u * kBlock) * ldw] = nv;
                vals[rr + u * kBlock] = static_cast<double>(nv);
            }
        }
        for (; rr < nr; rr += kBlock) {
            const int64_t i = r0 + rr;
            double acc = 0.0;
            for (int o = 0; o < k; ++o) acc = fma(static_cast<double>(B[i * ldb + o]), cs[o], acc);
            const T nv = static_cast<T>(static_cast<double>(w[i * ldw]) - acc);
            w[i * ldw] = nv;
            vals[rr] = static_cast<double>(nv);
        }
    } else {
        // four rows per wavefront trip: their loads and reductions interleave
        const int lane = t % kWave, wv = t / kWave;
        int rr = wv;
        for (; rr + 3 * dev::kWaves < nr; rr += 4 * dev::kWaves) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int o = lane; o < k; o += kWave) {
                const double co = c[o];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[u] = fma(static_cast<double>(B[(r0 + rr + u * dev::kWaves) * ldb + o]), co, acc[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] = dev::wave_sum(acc[u]);
            if (lane == 0) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t i = r0 + rr + u * dev::kWaves;
                    const T nv = static_cast<T>(static_cast<double>(w[i * ldw]) - acc[u]);
                    w[i * ldw] = nv;
                    vals[rr + u * dev::kWaves] = static_cast<double>(nv);
                }
            }
        }
        for (; rr < nr; rr += dev::kWaves) {
            const int64_t i = r0 + rr;
            double acc = 0.0;
            for (int o = lane; o < k; o += kWave) acc = fma(static_cast<double>(B[i * ldb + o]), c[o], acc);
            acc = dev::wave_sum(acc);
            if (lane == 0) {
                const T nv = static_cast<T>(static_cast<double>(w[i * ldw]) - acc);
                w[i * ldw] = nv;
                vals[rr] = static_cast<double>(nv);
            }
        }
    }
    __syncthreads();
    double bm, ss, bb;
    dev::block_stats(vals, nr, sm, bm, ss, bb);
    if (!dev::publish_stats(bm, ss, bb, partial, out, counter, sm)) return;
    if (gate_out != nullptr && t == 0) {
        const double na = dev::nominal_norm(out), nb = dev::nominal_norm(before);
        *gate_out = (isfinite(nb) && isfinite(na) && na < factor * nb) ? 1.0 : 0.0;
    }
}

}  // namespace

template <typename TB, typename TW>
void rank_one_project(hipStream_t s, int64_t N, int k, const TB* B, int64_t ldb, const TW* w, int64_t ldw,
                      double* c, double* coeff, bool accumulate, Scratch& scratch, unsigned* counter,
                      const double* gate_a, const double* gate_b) {
    if (N <= 0 || k <= 0) return;
    if (k > kBlock) {
        // row chunks of at least 64 rows, at most 64 of them
        const int64_t nbr = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(N, 64)));
        const int64_t rpb = ceil_div(N, nbr);
        const int64_t chunks = ceil_div(N, rpb);
        scratch.ensure(static_cast<size_t>(chunks * k) * sizeof(double));
        double* partial = static_cast<double*>(scratch.ptr);
        const dim3 grid(static_cast<unsigned>(ceil_div(k, kBlock)), static_cast<unsigned>(chunks));
        hipLaunchKernelGGL((project_wide_kernel<TB, TW>), grid, dim3(kBlock), 0, s, N, k, B, ldb, w, ldw, rpb,
                           partial, gate_a, gate_b);
        AMFLS_LAUNCH_CHECK();
        hipLaunchKernelGGL(project_combine_kernel, dim3(static_cast<unsigned>(ceil_div(k, kCombineCols))),
                           dim3(kBlock), 0, s, static_cast<int>(chunks), k, static_cast<const double*>(partial), c,
                           coeff, accumulate ? 1 : 0, gate_a, gate_b);
        AMFLS_LAUNCH_CHECK();
        return;
    }
    const int64_t work = N * static_cast<int64_t>(k);
    int64_t nb = std::max<int64_t>(ceil_div(N, 1024), ceil_div(work, 16384));
    nb = std::max<int64_t>(1, std::min<int64_t>(nb, 256));
    const int64_t rpb = ceil_div(N, nb);
    nb = ceil_div(N, rpb);
    scratch.ensure(static_cast<size_t>(nb * k) * sizeof(double));
    hipLaunchKernelGGL((project_kernel<TB, TW>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, N, k, B, ldb,
                       w, ldw, rpb, static_cast<double*>(scratch.ptr), c, coeff, accumulate ? 1 : 0, counter,
                       gate_a, gate_b);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void rank_one_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* c, T* w,
                     int64_t ldw, double* out, Scratch& scratch, unsigned* counter, const double* gate_a,
                     const double* gate_b, const double* before, double factor, double* gate_out) {
    if (N <= 0) return;
    // about 256 blocks of at least 64 rows (smaller blocks, more of them, made
    // the wide-basis update slower: every block takes a ticket on one counter)
    const int rpb = static_cast<int>(std::max<int64_t>(64, std::min<int64_t>(1024, ceil_div(N, 256))));
    const int64_t nb = ceil_div(N, rpb);
    scratch.ensure(static_cast<size_t>(3 * nb) * sizeof(double));
    const size_t lds = static_cast<size_t>(rpb) * sizeof(double);
    double* partial = static_cast<double*>(scratch.ptr);
    if (k <= 16)
        hipLaunchKernelGGL((update_kernel<T, true>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), lds, s, N, k, B,
                           ldb, c, w, ldw, rpb, partial, out, counter, gate_a, gate_b, before, factor, gate_out);
    else
        hipLaunchKernelGGL((update_kernel<T, false>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), lds, s, N, k,
                           B, ldb, c, w, ldw, rpb, partial, out, counter, gate_a, gate_b, before, factor, gate_out);
    AMFLS_LAUNCH_CHECK();
}

template void rank_one_project<double, double>(hipStream_t, int64_t, int, const double*, int64_t, const double*,
                                               int64_t, double*, double*, bool, Scratch&, unsigned*,
                                               const double*, const double*);
template void rank_one_project<float, float>(hipStream_t, int64_t, int, const float*, int64_t, const float*,
                                             int64_t, double*, double*, bool, Scratch&, unsigned*,
                                             const double*, const double*);
template void rank_one_project<float, double>(hipStream_t, int64_t, int, const float*, int64_t, const double*,
                                              int64_t, double*, double*, bool, Scratch&, unsigned*,
                                              const double*, const double*);
template void rank_one_update<double>(hipStream_t, int64_t, int, const double*, int64_t, const double*, double*,
                                      int64_t, double*, Scratch&, unsigned*, const double*, const double*,
                                      const double*, double, double*);
template void rank_one_update<float>(hipStream_t, int64_t, int, const float*, int64_t, const double*, float*,
                                     int64_t, double*, Scratch&, unsigned*, const double*, const double*,
                                     const double*, double, double*);

}  // namespace amflsx
