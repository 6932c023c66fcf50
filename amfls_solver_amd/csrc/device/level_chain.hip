// Kernels of the batched rank-one Golub-Kahan level (gfx950).
//
// A one-column level is a chain of tall-vector passes over a residual y
// against a basis B (N x k, row-major): the coupling update y = r - a p, the
// projection c = B^T y, the update y -= B c, a possible second (DGKS) pass
// and the normalization.  The chain runs in four kernels per side:
//
//   couple_project  y = r - a p, stats(r), stats(y), partials of B^T y
//   update_project  y -= B c1, stats(y), partials of B^T y (for pass two)
//   update_second   only if ||y|| < ||y_before|| / sqrt 2: y -= B c2, stats(y)
//   normalize       rank rule, dst = y / ||y||, partials of dst^T b
//
// Every reduction is split between its producer, which writes one partial
// per block, and its consumer kernel, whose every block combines all
// partials in the same fixed order and so obtains the same value.  Kernel
// boundaries are the only grid-wide synchronization: no tickets, no device-
// scope fences, no extra launches for final reductions.  Scalars live in
// device memory, so whole batches of levels are enqueued without the host.
//
// Norm statistics are accumulated in one pass: each thread keeps its max
// |v|, the frexp exponent e of that max and sum (|v| 2^-e)^2, rescaling the
// sum by an exact power of two when the max grows; blocks and the consumer
// combine these triples the same way.  The result equals the two-pass
// statistics up to summation order.
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace amflsx {
namespace {

using dev::kBlock;

struct Stat {
    double m = 0.0;    // max |v| over finite values
    int e = 0;         // frexp exponent of m (0 while m == 0)
    double s = 0.0;    // sum (|v| 2^-e)^2
    double bad = 0.0;  // nonfinite count
    __device__ void add(double v) {
        v = fabs(v);
        if (!isfinite(v)) {
            bad += 1.0;
            return;
        }
        if (v > m) {
            int ev = 0;
            (void)frexp(v, &ev);
            if (ev > e && m > 0.0) s = ldexp(s, 2 * (e - ev));
            if (m == 0.0 || ev > e) e = ev;
            m = v;
        }
        if (v > 0.0) {
            const double sv = ldexp(v, -e);
            s = fma(sv, sv, s);
        }
    }
};

// Wave combine of per-lane Stat (max, then sums rescaled exactly to the wave
// max exponent, butterfly order); lane 0 writes one triple per wavefront at
// part[3 * (block * kWaves + wave)].  No barriers.
__device__ void wave_publish(const Stat& st, double* part) {
    const double wm = dev::wave_max(st.m);
    int ew = 0;
    if (wm > 0.0) (void)frexp(wm, &ew);
    const double ss = dev::wave_sum(st.m > 0.0 ? ldexp(st.s, 2 * (st.e - ew)) : 0.0);
    const double bb = dev::wave_sum(st.bad);
    if (threadIdx.x % kWave == 0) {
        const int64_t slot = static_cast<int64_t>(blockIdx.x) * dev::kWaves + threadIdx.x / kWave;
        part[3 * slot] = wm;
        part[3 * slot + 1] = ss;
        part[3 * slot + 2] = bb;
    }
}

// Statistics triple of np per-wavefront triples, computed by every wavefront
// on its own (shuffles only, no barriers); every lane receives it.
__device__ void combine_stats(const double* part, int np, double out[3]) {
    const int lane = threadIdx.x % kWave;
    double gm = 0.0;
    for (int q = lane; q < np; q += kWave) gm = fmax(gm, part[3 * q]);
    gm = dev::wave_max(gm);
    int e = 0;
    if (gm > 0.0) (void)frexp(gm, &e);
    double tot = 0.0, bad = 0.0;
    for (int q = lane; q < np; q += kWave) {
        const double pm = part[3 * q];
        bad += part[3 * q + 2];
        if (pm > 0.0) {
            int eb = 0;
            (void)frexp(pm, &eb);
            tot += ldexp(part[3 * q + 1], 2 * (eb - e));
        }
    }
    out[0] = gm;
    out[1] = dev::wave_sum(tot);
    out[2] = dev::wave_sum(bad);
}

// DGKS second-pass test from the partials of ||before|| and ||after||.
__device__ bool second_pass_needed(const double* before, int np, const double* after, double factor) {
    double sb[3], sa[3];
    combine_stats(before, np, sb);
    combine_stats(after, np, sa);
    const double nb = dev::nominal_norm(sb), na = dev::nominal_norm(sa);
    return isfinite(nb) && isfinite(na) && na < factor * nb;
}

// c[o] = sum_b Q[b][o] in a fixed order, into shared memory (k values): one
// wavefront per coefficient (lanes over blocks, butterfly) for k <= 32, one
// thread per coefficient (blocks in order) otherwise.
__device__ void combine_coefficients(const double* Q, int nbq, int k, double* cs) {
    if (k <= 32) {
        const int lane = threadIdx.x % kWave;
        for (int o = threadIdx.x / kWave; o < k; o += dev::kWaves) {
            double acc = 0.0;
            for (int b = lane; b < nbq; b += kWave) acc += Q[static_cast<int64_t>(b) * k + o];
            acc = dev::wave_sum(acc);
            if (lane == 0) cs[o] = acc;
        }
        __syncthreads();
        return;
    }
    for (int o = threadIdx.x; o < k; o += kBlock) {
        double acc = 0.0;
        int b = 0;
        for (; b + 3 < nbq; b += 4) {
            double pv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pv[u] = Q[static_cast<int64_t>(b + u) * k + o];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += pv[u];
        }
        for (; b < nbq; ++b) acc += Q[static_cast<int64_t>(b) * k + o];
        cs[o] = acc;
    }
    __syncthreads();
}

// Per-block projection partials Q[b][o] = sum over the block's rows of
// B[i, o] * y[i], y read from shared memory (ys[row - r0]).
template <typename T>
__device__ void project_block(const T* __restrict__ B, int64_t ldb, int k, int64_t r0, int nr, const double* ys,
                              double* red, double* Q) {
    const int t = threadIdx.x;
    for (int o0 = 0; o0 < k; o0 += kBlock) {
        const int kc = std::min(kBlock, k - o0);
        int kp = 1;
        while (kp < kc) kp <<= 1;
        const int S = kBlock / kp;
        const int o = t % kp, sidx = t / kp;
        double acc = 0.0;
        if (o < kc) {
            const int oc = o0 + o;
            int i = sidx;
            for (; i + 7 * S < nr; i += 8 * S) {
                double bv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) bv[u] = static_cast<double>(B[(r0 + i + u * S) * ldb + oc]);
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = fma(bv[u], ys[i + u * S], acc);
            }
            for (; i < nr; i += S) acc = fma(static_cast<double>(B[(r0 + i) * ldb + oc]), ys[i], acc);
        }
        red[t] = acc;
        __syncthreads();
        for (int stride = S / 2; stride > 0; stride >>= 1) {
            if (sidx < stride) red[t] += red[t + stride * kp];
            __syncthreads();
        }
        if (t < kc) Q[static_cast<int64_t>(blockIdx.x) * k + o0 + t] = red[t];
        __syncthreads();
    }
}

// y = r - (*coef) p (in place on r), stats of r and y, projection partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void couple_project_kernel(int64_t N, int k, int R, T* r, int64_t ldr,
                                                                const T* p, int64_t ldp, const double* coef,
                                                                const T* __restrict__ B, int64_t ldb,
                                                                const double* go, double* part_r, double* part_y,
                                                                double* Q) {
    if (*go == 0.0) return;
    extern __shared__ double ys[];  // R values of y
    __shared__ double red[kBlock];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    const double a = *coef;
    Stat sr, sy;
    for (int i = threadIdx.x; i < nr; i += kBlock) {
        const int64_t g = r0 + i;
        const double rv = static_cast<double>(r[g * ldr]);
        const T yv = static_cast<T>(rv - static_cast<double>(p[g * ldp]) * a);
        r[g * ldr] = yv;
        ys[i] = static_cast<double>(yv);
        sr.add(rv);
        sy.add(static_cast<double>(yv));
    }
    wave_publish(sr, part_r);
    wave_publish(sy, part_y);
    __syncthreads();
    project_block<T>(B, ldb, k, r0, nr, ys, red, Q);
}

// y -= B c with c = sum_b Qin[b] (gated: second pass only when the DGKS
// test on (before, after) holds); stats of the new y; with Qout, projection
// partials of the new y for a following pass.
template <typename T>
__global__ __launch_bounds__(kBlock) void update_kernel(int64_t N, int k, int R, const T* __restrict__ B,
                                                        int64_t ldb, const double* Qin, int nbq, T* y, int64_t ldy,
                                                        double* part_y, double* Qout, const double* go,
                                                        const double* before, const double* after, int np_stats,
                                                        double factor) {
    extern __shared__ double lds[];  // c (k values) then the block's new y (R values)
    double* cs = lds;
    double* ys = lds + k;
    __shared__ double red[kBlock];
    if (*go == 0.0) return;
    if (before != nullptr && !second_pass_needed(before, np_stats, after, factor)) return;
    combine_coefficients(Qin, nbq, k, cs);
    const int t = threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    Stat st;
    if (k <= 16) {
        for (int rr = t; rr < nr; rr += kBlock) {
            const int64_t i = r0 + rr;
            double acc = 0.0;
            for (int o = 0; o < k; ++o) acc = fma(static_cast<double>(B[i * ldb + o]), cs[o], acc);
            const T nv = static_cast<T>(static_cast<double>(y[i * ldy]) - acc);
            y[i * ldy] = nv;
            ys[rr] = static_cast<double>(nv);
            st.add(static_cast<double>(nv));
        }
    } else {
        const int lane = t % kWave, wv = t / kWave;
        int rr = wv;
        for (; rr + 3 * dev::kWaves < nr; rr += 4 * dev::kWaves) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int o = lane; o < k; o += kWave) {
                const double co = cs[o];
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
                    const T nv = static_cast<T>(static_cast<double>(y[i * ldy]) - acc[u]);
                    y[i * ldy] = nv;
                    ys[rr + u * dev::kWaves] = static_cast<double>(nv);
                    st.add(static_cast<double>(nv));
                }
            }
        }
        for (; rr < nr; rr += dev::kWaves) {
            const int64_t i = r0 + rr;
            double acc = 0.0;
            for (int o = lane; o < k; o += kWave) acc = fma(static_cast<double>(B[i * ldb + o]), cs[o], acc);
            acc = dev::wave_sum(acc);
            if (lane == 0) {
                const T nv = static_cast<T>(static_cast<double>(y[i * ldy]) - acc);
                y[i * ldy] = nv;
                ys[rr] = static_cast<double>(nv);
                st.add(static_cast<double>(nv));
            }
        }
    }
    wave_publish(st, part_y);
    if (Qout != nullptr) {
        __syncthreads();
        project_block<T>(B, ldb, k, r0, nr, ys, red, Qout);
    }
}

// Rank decision and normalization of y (the host rule for a one-column
// block): ||scale|| from part_scale, ||resid|| from the first-pass partials,
// or from the second-pass ones when the DGKS test on (before, first) held.
// Accepted when both are finite and ||resid|| > rank_factor * ||scale||;
// then dst = y * (1 / ||resid||) and, with b given, gpart[block] = sum over
// the block of dst[i] b[i].  Block 0 writes record = {resid, scale,
// accepted, c} (c = summed coefficient of a one-column basis, when Q1 is
// given) and, on acceptance, *cell = resid; on rejection *go = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void chain_normalize_kernel(
    int64_t N, int R, const T* y, int64_t ldy, T* dst, int64_t ldd, const double* part_scale,
    const double* part_before, const double* part_first, const double* part_second, int np_stats,
    double dgks_factor, double rank_factor, const double* Q1, const double* Q2, int nbq, const double* b,
    double* gpart, double* record, double* cell, double* go) {
    __shared__ double sm[dev::kWaves];
    if (*go == 0.0) return;
    double ss[3], sf[3];
    combine_stats(part_scale, np_stats, ss);
    const bool second = second_pass_needed(part_before, np_stats, part_first, dgks_factor);
    combine_stats(second ? part_second : part_first, np_stats, sf);
    const double scale = dev::nominal_norm(ss), resid = dev::nominal_norm(sf);
    const bool ok = isfinite(scale) && scale >= 0.0 && isfinite(resid) && resid > rank_factor * scale;
    double c_total = 0.0;
    if (Q1 != nullptr && blockIdx.x == 0 && threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        double c1 = 0.0, c2 = 0.0;
        for (int q = lane; q < nbq; q += kWave) c1 += Q1[q];
        c1 = dev::wave_sum(c1);
        if (second) {
            for (int q = lane; q < nbq; q += kWave) c2 += Q2[q];
            c2 = dev::wave_sum(c2);
        }
        c_total = second ? c1 + c2 : c1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        record[0] = resid;
        record[1] = scale;
        record[2] = ok ? 1.0 : 0.0;
        if (Q1 != nullptr) {
            // the coefficient sums of a one-column basis, in the order the
            // update kernels combined them (wave butterfly over blocks)
            record[3] = c_total;
        }
        if (ok) *cell = resid;
        else *go = 0.0;
    }
    if (!ok) return;
    const double inv = 1.0 / resid;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    double acc = 0.0;
    for (int i = threadIdx.x; i < nr; i += kBlock) {
        const int64_t g = r0 + i;
        const double v = static_cast<double>(y[g * ldy]);
        const T nv = static_cast<T>(inv == 1.0 ? v : inv * v);
        dst[g * ldd] = nv;
        if (b != nullptr) acc = fma(static_cast<double>(nv), b[g], acc);
    }
    if (b != nullptr) {
        acc = dev::block_sum(acc, sm);
        if (threadIdx.x == 0) gpart[blockIdx.x] = acc;
    }
}

}  // namespace

int chain_block_rows(int64_t N, int k) {
    // about 128 blocks; fewer, longer blocks when the k projection partials
    // each consumer block combines would exceed 16K values
    const int64_t base = std::max<int64_t>(64, ceil_div(N, 128));
    const int64_t wide = ceil_div(N * std::max(k, 1), 16384);
    return static_cast<int>(std::min<int64_t>(4096, std::max(base, wide)));
}

template <typename T>
void chain_couple_project(hipStream_t s, int64_t N, int k, T* r, int64_t ldr, const T* p, int64_t ldp,
                          const double* coef, const T* B, int64_t ldb, const double* go, double* part_r,
                          double* part_y, double* Q) {
    const int R = chain_block_rows(N, k);
    const int64_t nb = ceil_div(N, R);
    hipLaunchKernelGGL(couple_project_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), R * sizeof(double), s,
                       N, k, R, r, ldr, p, ldp, coef, B, ldb, go, part_r, part_y, Q);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void chain_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* Qin, T* y, int64_t ldy,
                  double* part_y, double* Qout, const double* go, const double* before, const double* after,
                  double factor) {
    const int R = chain_block_rows(N, k);
    const int64_t nb = ceil_div(N, R);
    const size_t lds = static_cast<size_t>(k + R) * sizeof(double);
    hipLaunchKernelGGL(update_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), lds, s, N, k, R, B, ldb, Qin,
                       static_cast<int>(nb), y, ldy, part_y, Qout, go, before, after,
                       static_cast<int>(nb * dev::kWaves), factor);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void chain_normalize(hipStream_t s, int64_t N, int k, const T* y, int64_t ldy, T* dst, int64_t ldd,
                     const double* part_scale, const double* part_before, const double* part_first,
                     const double* part_second, double dgks_factor, double rank_factor, const double* Q1,
                     const double* Q2, const double* b, double* gpart, double* record, double* cell, double* go) {
    const int R = chain_block_rows(N, k);
    const int64_t nb = ceil_div(N, R);
    hipLaunchKernelGGL(chain_normalize_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, N, R, y, ldy,
                       dst, ldd, part_scale, part_before, part_first, part_second,
                       static_cast<int>(nb * dev::kWaves), dgks_factor,
                       rank_factor, Q1, Q2, static_cast<int>(nb), b, gpart, record, cell, go);
    AMFLS_LAUNCH_CHECK();
}

#define AMFLS_CHAIN(T)                                                                                             \
    template void chain_couple_project<T>(hipStream_t, int64_t, int, T*, int64_t, const T*, int64_t,             \
                                          const double*, const T*, int64_t, const double*, double*, double*,     \
                                          double*);                                                              \
    template void chain_update<T>(hipStream_t, int64_t, int, const T*, int64_t, const double*, T*, int64_t,       \
                                  double*, double*, const double*, const double*, const double*, double);       \
    template void chain_normalize<T>(hipStream_t, int64_t, int, const T*, int64_t, T*, int64_t, const double*,   \
                                     const double*, const double*, const double*, double, double, const double*, \
                                     const double*, const double*, double*, double*, double*, double*);
AMFLS_CHAIN(double)
AMFLS_CHAIN(float)
#undef AMFLS_CHAIN

}  // namespace amflsx
