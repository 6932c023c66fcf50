// Device-side reduction helpers shared by the AMFLS kernels (wave64,
// 256-thread blocks): wavefront and block reductions, the one-launch grid
// reduction of norm statistics (per-block partials plus a last-block
// combine), and the nominal norm of a statistics triple.
#pragma once

#include "common.hpp"

namespace amflsx {
namespace dev {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

// A gate is an optional device flag: work is skipped while it holds 0.
__device__ __forceinline__ bool gated_off(const double* gate) { return gate != nullptr && *gate == 0.0; }

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

// Sum over the S = kBlock / kp threads t = o + kp * sidx that share output o
// (kp a power of two): within a wavefront by a butterfly over the sidx bits,
// then the kWaves wavefront sums in order (kp < kWave); with kp >= kWave the
// S <= 4 partial sums of o lie in different wavefronts and are added in
// order.  One barrier either way (two with the reuse guard); red holds kBlock
// doubles.  Returns the sum in threads t < kp.
__device__ __forceinline__ double reduce_shared_outputs(double acc, int kp, double* red) {
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    double v = 0.0;
    if (kp < kWave) {
        for (int off = kp; off < kWave; off <<= 1) acc += __shfl_xor(acc, off);
        if (lane < kp) red[w * kp + lane] = acc;
        __syncthreads();
        if (t < kp) {
#pragma unroll
            for (int w2 = 0; w2 < kWaves; ++w2) v += red[w2 * kp + t];
        }
    } else {
        red[t] = acc;
        __syncthreads();
        if (t < kp) {
            for (int q = 0; q < kBlock / kp; ++q) v += red[t + q * kp];
        }
    }
    __syncthreads();
    return v;
}

// ||v|| from (max |v|, sum (|v| 2^-e)^2, nonfinite count), e = frexp
// exponent of the max; NaN when a value was nonfinite.
__device__ __forceinline__ double nominal_norm(const double* st) {
    if (st[2] > 0.0) return __builtin_nan("");
    if (st[0] == 0.0) return 0.0;
    int e = 0;
    (void)frexp(st[0], &e);
    return ldexp(sqrt(st[1]), e);
}

// Block-wide max / sum; every thread receives the result.  `sm` holds
// kWaves doubles of shared memory; the call ends with the block synchronized.
__device__ __forceinline__ double block_max(double v, double* sm) {
    v = wave_max(v);
    __syncthreads();
    if (threadIdx.x % kWave == 0) sm[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = sm[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r = fmax(r, sm[k]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_sum(double v, double* sm) {
    v = wave_sum(v);
    __syncthreads();
    if (threadIdx.x % kWave == 0) sm[threadIdx.x / kWave] = v;
    __syncthreads();
    double r = sm[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r += sm[k];
    __syncthreads();
    return r;
}

// Statistics triple of the n values in shared memory `vals` (max |v| over
// finite values, sum (|v| 2^-eb)^2 with eb the frexp exponent of that max,
// nonfinite count), returned to every thread of the block.
__device__ __forceinline__ void block_stats(const double* vals, int n, double* sm, double& bm, double& ss,
                                            double& bb) {
    double m = 0.0, bad = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const double v = fabs(vals[i]);
        if (isfinite(v)) m = fmax(m, v); else bad += 1.0;
    }
    bm = block_max(m, sm);
    bb = block_sum(bad, sm);
    int eb = 0;
    if (bm > 0.0) (void)frexp(bm, &eb);
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const double v = fabs(vals[i]);
        if (isfinite(v)) {
            const double sv = ldexp(v, -eb);
            acc = fma(sv, sv, acc);
        }
    }
    ss = block_sum(acc, sm);
}

// Grid combine of per-block statistics triples.  Every block calls it with
// its triple; the last block to take a ticket on `counter` rescales each
// partial sum exactly (by 2^(2(eb - e)), e the global max exponent), writes
// out[0..2], resets the counter and returns true (all threads); other blocks
// return false.
__device__ __forceinline__ bool publish_stats(double bm, double ss, double bb, double* partial, double* out,
                                              unsigned* counter, double* sm) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = bm;
        partial[3 * blockIdx.x + 1] = ss;
        partial[3 * blockIdx.x + 2] = bb;
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    const int nb = static_cast<int>(gridDim.x);
    double gm = 0.0;
    // plain loads: the acquire fence above made every published triple visible
    for (int b = threadIdx.x; b < nb; b += kBlock) gm = fmax(gm, partial[3 * b]);
    gm = block_max(gm, sm);
    int e = 0;
    if (gm > 0.0) (void)frexp(gm, &e);
    double tot = 0.0, tbad = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        const double pm = partial[3 * b];
        const double ps = partial[3 * b + 1];
        tbad += partial[3 * b + 2];
        if (pm > 0.0) {
            int eb = 0;
            (void)frexp(pm, &eb);
            tot += ldexp(ps, 2 * (eb - e));
        }
    }
    tot = block_sum(tot, sm);
    tbad = block_sum(tbad, sm);
    if (threadIdx.x == 0) {
        out[0] = gm;
        out[1] = tot;
        out[2] = tbad;
        *counter = 0u;
    }
    __syncthreads();
    return true;
}

}  // namespace dev
}  // namespace amflsx
