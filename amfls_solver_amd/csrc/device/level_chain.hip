// Kernels of the batched rank-one Golub-Kahan level (gfx950).
//
// A one-column level is a chain of tall-vector passes: the product, the
// coupling update y = r - a p, the projection c = B^T y, the update
// y -= B c, a possible second (DGKS) pass, and the normalization.  Every
// reduction in the chain is split between its producer, which writes one
// partial per block, and its consumer kernel, whose every block combines
// all partials in the same fixed order and so obtains the same value.
// Kernel boundaries are the only grid-wide synchronization: no tickets, no
// device-scope fences, no extra launches for final reductions.  Scalars
// (coupling coefficients, the continue flag) live in device memory, so a
// whole batch of levels is enqueued without host involvement.
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace amflsx {
namespace {

using dev::kBlock;

// Statistics triple of nb per-block triples (max, scaled square sum with
// the block's own max exponent, nonfinite count), rescaled exactly to the
// global max exponent; every thread receives it.
__device__ void combine_stats(const double* part, int nb, double* sm, double out[3]) {
    double gm = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) gm = fmax(gm, part[3 * b]);
    gm = dev::block_max(gm, sm);
    int e = 0;
    if (gm > 0.0) (void)frexp(gm, &e);
    double tot = 0.0, bad = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        const double pm = part[3 * b];
        bad += part[3 * b + 2];
        if (pm > 0.0) {
            int eb = 0;
            (void)frexp(pm, &eb);
            tot += ldexp(part[3 * b + 1], 2 * (eb - e));
        }
    }
    out[0] = gm;
    out[1] = dev::block_sum(tot, sm);
    out[2] = dev::block_sum(bad, sm);
}

// Block statistics of n values in shared memory, written as one triple.
__device__ void publish_block(const double* vals, int n, double* sm, double* part) {
    double bm, ss, bb;
    dev::block_stats(vals, n, sm, bm, ss, bb);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = bm;
        part[3 * blockIdx.x + 1] = ss;
        part[3 * blockIdx.x + 2] = bb;
    }
}

// DGKS second-pass test from the partials of ||before|| and ||after||.
__device__ bool second_pass_needed(const double* before, int nb_before, const double* after, int nb_after,
                                   double factor, double* sm) {
    double sb[3], sa[3];
    combine_stats(before, nb_before, sm, sb);
    combine_stats(after, nb_after, sm, sa);
    const double nb = dev::nominal_norm(sb), na = dev::nominal_norm(sa);
    return isfinite(nb) && isfinite(na) && na < factor * nb;
}

// Optional second-pass gate shared by the projection and update kernels.
struct PassGate {
    const double* before;  // null: first pass, always runs
    const double* after;
    int nb_before, nb_after;
    double factor;
};

__device__ bool pass_skipped(const double* go, const PassGate& g, double* sm) {
    if (*go == 0.0) return true;
    if (g.before == nullptr) return false;
    return !second_pass_needed(g.before, g.nb_before, g.after, g.nb_after, g.factor, sm);
}

// y = r - (*coef) * p over rows [b R, (b+1) R), in place on r; per-block
// statistics of r (before the update) and of y.
template <typename T>
__global__ __launch_bounds__(kBlock) void couple_kernel(int64_t N, int R, T* r, int64_t ldr, const T* p, int64_t ldp,
                                                        const double* coef, const double* go, double* part_r,
                                                        double* part_y) {
    if (*go == 0.0) return;
    extern __shared__ double lds[];
    double* vr = lds;
    double* vy = lds + R;
    __shared__ double sm[dev::kWaves];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    const double a = *coef;
    for (int i = threadIdx.x; i < nr; i += kBlock) {
        const int64_t g = r0 + i;
        const double rv = static_cast<double>(r[g * ldr]);
        const T yv = static_cast<T>(rv - static_cast<double>(p[g * ldp]) * a);
        r[g * ldr] = yv;
        vr[i] = rv;
        vy[i] = static_cast<double>(yv);
    }
    __syncthreads();
    publish_block(vr, nr, sm, part_r);
    publish_block(vy, nr, sm, part_y);
}

// Partial projections P[b][o] = sum over the block's rows of B[i, o] y[i].
template <typename T>
__global__ __launch_bounds__(kBlock) void chain_project_kernel(int64_t N, int k, int R, const T* __restrict__ B,
                                                               int64_t ldb, const T* __restrict__ y, int64_t ldy,
                                                               double* __restrict__ P, const double* go,
                                                               PassGate gate) {
    __shared__ double sm[dev::kWaves];
    __shared__ double red[kBlock];
    if (pass_skipped(go, gate, sm)) return;
    const int t = threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int64_t r1 = std::min<int64_t>(N, r0 + R);
    for (int o0 = 0; o0 < k; o0 += kBlock) {
        const int kc = std::min(kBlock, k - o0);
        int kp = 1;
        while (kp < kc) kp <<= 1;
        const int S = kBlock / kp;
        const int o = t % kp, sidx = t / kp;
        double acc = 0.0;
        if (o < kc) {
            const int oc = o0 + o;
            int64_t i = r0 + sidx;
            for (; i + 7 * S < r1; i += 8 * S) {
                double bv[8], wv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    bv[u] = static_cast<double>(B[(i + u * S) * ldb + oc]);
                    wv[u] = static_cast<double>(y[(i + u * S) * ldy]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = fma(bv[u], wv[u], acc);
            }
            for (; i < r1; i += S) acc = fma(static_cast<double>(B[i * ldb + oc]), static_cast<double>(y[i * ldy]), acc);
        }
        red[t] = acc;
        __syncthreads();
        for (int stride = S / 2; stride > 0; stride >>= 1) {
            if (sidx < stride) red[t] += red[t + stride * kp];
            __syncthreads();
        }
        if (t < kc) P[static_cast<int64_t>(blockIdx.x) * k + o0 + t] = red[t];
        __syncthreads();
    }
}

// c = sum_b P[b][:] (block order), y -= B c, per-block statistics of y.
template <typename T>
__global__ __launch_bounds__(kBlock) void chain_update_kernel(int64_t N, int k, int R, const T* __restrict__ B,
                                                              int64_t ldb, const double* __restrict__ P, int nbP,
                                                              T* y, int64_t ldy, double* part_y, const double* go,
                                                              PassGate gate) {
    extern __shared__ double lds[];  // c (k) then the block's new values (R)
    double* cs = lds;
    double* vals = lds + k;
    __shared__ double sm[dev::kWaves];
    if (pass_skipped(go, gate, sm)) return;
    const int t = threadIdx.x;
    for (int o = t; o < k; o += kBlock) {
        double acc = 0.0;
        int b = 0;
        for (; b + 3 < nbP; b += 4) {
            double pv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pv[u] = P[static_cast<int64_t>(b + u) * k + o];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += pv[u];
        }
        for (; b < nbP; ++b) acc += P[static_cast<int64_t>(b) * k + o];
        cs[o] = acc;
    }
    __syncthreads();
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    if (k <= 16) {
        for (int rr = t; rr < nr; rr += kBlock) {
            const int64_t i = r0 + rr;
            double acc = 0.0;
            for (int o = 0; o < k; ++o) acc = fma(static_cast<double>(B[i * ldb + o]), cs[o], acc);
            const T nv = static_cast<T>(static_cast<double>(y[i * ldy]) - acc);
            y[i * ldy] = nv;
            vals[rr] = static_cast<double>(nv);
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
                    vals[rr + u * dev::kWaves] = static_cast<double>(nv);
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
                vals[rr] = static_cast<double>(nv);
            }
        }
    }
    __syncthreads();
    publish_block(vals, nr, sm, part_y);
}

// Rank decision and normalization of the residual y (the host rule for a
// one-column block): ||scale|| from the producer's partials, ||resid|| from
// the first-pass partials, or from the second-pass ones when the DGKS test
// on (before, first) asked for that pass.  Accepted when both are finite
// and ||resid|| > rank_factor * ||scale||; then dst = y * (1 / ||resid||)
// and, with b given, gpart[block] = sum over the block of dst[i] b[i].
// Block 0 writes record = {resid, scale, accepted, c} (c = the summed
// projection coefficient of a one-column basis, when P1 is given) and, on
// acceptance, *cell = resid; on rejection *go = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void chain_normalize_kernel(
    int64_t N, int R, const T* y, int64_t ldy, T* dst, int64_t ldd, const double* part_scale, int nb_scale,
    const double* part_before, const double* part_first, const double* part_second, int nb_update,
    double dgks_factor, double rank_factor, const double* P1, const double* P2, int nbP, const double* b,
    double* gpart, double* record, double* cell, double* go) {
    __shared__ double sm[dev::kWaves];
    __shared__ double red[kBlock];
    if (*go == 0.0) return;
    double ss[3], sf[3];
    combine_stats(part_scale, nb_scale, sm, ss);
    const bool second = second_pass_needed(part_before, nb_scale, part_first, nb_update, dgks_factor, sm);
    combine_stats(second ? part_second : part_first, nb_update, sm, sf);
    const double scale = dev::nominal_norm(ss), resid = dev::nominal_norm(sf);
    const bool ok = isfinite(scale) && scale >= 0.0 && isfinite(resid) && resid > rank_factor * scale;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        record[0] = resid;
        record[1] = scale;
        record[2] = ok ? 1.0 : 0.0;
        if (P1 != nullptr) {
            double c = 0.0;
            for (int q = 0; q < nbP; ++q) c += P1[q];
            if (second) {
                double c2 = 0.0;
                for (int q = 0; q < nbP; ++q) c2 += P2[q];
                c += c2;
            }
            record[3] = c;
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
    (void)red;
}

}  // namespace

int chain_rows_per_block(int64_t N) {
    return static_cast<int>(std::max<int64_t>(64, std::min<int64_t>(1024, ceil_div(N, 128))));
}

int chain_project_rows(int64_t N, int k) {
    const int64_t nb_max = std::max<int64_t>(1, std::min<int64_t>(128, 65536 / std::max(k, 1)));
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(nb_max, ceil_div(N * std::max(k, 1), 8192)));
    return static_cast<int>(ceil_div(N, nb));
}

template <typename T>
void chain_couple(hipStream_t s, int64_t N, T* r, int64_t ldr, const T* p, int64_t ldp, const double* coef,
                  const double* go, double* part_r, double* part_y) {
    const int R = chain_rows_per_block(N);
    const int64_t nb = ceil_div(N, R);
    hipLaunchKernelGGL(couple_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 2 * R * sizeof(double), s, N,
                       R, r, ldr, p, ldp, coef, go, part_r, part_y);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void chain_project(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const T* y, int64_t ldy, double* P,
                   const double* go, const double* before, const double* after, int nb_before, int nb_after,
                   double factor) {
    const int R = chain_project_rows(N, k);
    const int64_t nb = ceil_div(N, R);
    const PassGate g{before, after, nb_before, nb_after, factor};
    hipLaunchKernelGGL(chain_project_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, N, k, R, B, ldb,
                       y, ldy, P, go, g);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void chain_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* P, int nbP, T* y,
                  int64_t ldy, double* part_y, const double* go, const double* before, const double* after,
                  int nb_before, int nb_after, double factor) {
    const int R = chain_rows_per_block(N);
    const int64_t nb = ceil_div(N, R);
    const PassGate g{before, after, nb_before, nb_after, factor};
    const size_t lds = static_cast<size_t>(k + R) * sizeof(double);
    hipLaunchKernelGGL(chain_update_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), lds, s, N, k, R, B, ldb,
                       P, nbP, y, ldy, part_y, go, g);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void chain_normalize(hipStream_t s, int64_t N, const T* y, int64_t ldy, T* dst, int64_t ldd, const double* part_scale,
                     int nb_scale, const double* part_before, const double* part_first, const double* part_second,
                     int nb_update, double dgks_factor, double rank_factor, const double* P1, const double* P2,
                     int nbP, const double* b, double* gpart, double* record, double* cell, double* go) {
    const int R = chain_rows_per_block(N);
    const int64_t nb = ceil_div(N, R);
    hipLaunchKernelGGL(chain_normalize_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, N, R, y, ldy,
                       dst, ldd, part_scale, nb_scale, part_before, part_first, part_second, nb_update, dgks_factor,
                       rank_factor, P1, P2, nbP, b, gpart, record, cell, go);
    AMFLS_LAUNCH_CHECK();
}

#define AMFLS_CHAIN(T)                                                                                            \
    template void chain_couple<T>(hipStream_t, int64_t, T*, int64_t, const T*, int64_t, const double*,          \
                                  const double*, double*, double*);                                            \
    template void chain_project<T>(hipStream_t, int64_t, int, const T*, int64_t, const T*, int64_t, double*,    \
                                   const double*, const double*, const double*, int, int, double);             \
    template void chain_update<T>(hipStream_t, int64_t, int, const T*, int64_t, const double*, int, T*, int64_t, \
                                  double*, const double*, const double*, const double*, int, int, double);     \
    template void chain_normalize<T>(hipStream_t, int64_t, const T*, int64_t, T*, int64_t, const double*, int,   \
                                     const double*, const double*, const double*, int, double, double,         \
                                     const double*, const double*, int, const double*, double*, double*,       \
                                     double*, double*);
AMFLS_CHAIN(double)
AMFLS_CHAIN(float)
#undef AMFLS_CHAIN

}  // namespace amflsx
