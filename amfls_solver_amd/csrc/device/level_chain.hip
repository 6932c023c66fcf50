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
// partials in the same fixed order and so obtains the same value.  The
// consumer issues the loads of all partials it needs together, so a kernel
// waits for at most a few dependent memory round trips; the update reads
// the basis once for both the update and the next projection.  Kernel
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
// rows per thread whose loads a row loop issues together (see the row loops)
constexpr int kTrip = 8;

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

// Merge a triple (pm = max, ps = sum scaled to pm's frexp exponent, pb =
// nonfinite count) into a running Stat, rescaling exactly by powers of two.
__device__ __forceinline__ void merge(Stat& st, double pm, double ps, double pb) {
    st.bad += pb;
    if (pm > 0.0) {
        int ep = 0;
        (void)frexp(pm, &ep);
        if (st.m == 0.0) {
            st.e = ep;
        } else if (ep > st.e) {
            st.s = ldexp(st.s, 2 * (st.e - ep));
            st.e = ep;
        }
        st.m = fmax(st.m, pm);
        st.s += ldexp(ps, 2 * (ep - st.e));
    }
}

// Wavefront combine of per-lane Stat (max, sums rescaled to the wave max
// exponent, butterfly order); every lane receives the triple in st.
__device__ __forceinline__ void wave_reduce(Stat& st) {
    const double wm = dev::wave_max(st.m);
    int ew = 0;
    if (wm > 0.0) (void)frexp(wm, &ew);
    const double ss = dev::wave_sum(st.m > 0.0 ? ldexp(st.s, 2 * (st.e - ew)) : 0.0);
    st.bad = dev::wave_sum(st.bad);
    st.m = wm;
    st.e = ew;
    st.s = ss;
}

// One statistics triple per block at part[3 * blockIdx.x]: wavefront
// combines, then the kWaves triples merged in wave order by thread 0.  All
// threads of the block must call it (it synchronizes); sm holds 3 * kWaves.
__device__ void block_publish(Stat st, double* part, double* sm) {
    wave_reduce(st);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        sm[3 * w] = st.m;
        sm[3 * w + 1] = st.s;
        sm[3 * w + 2] = st.bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Stat b;
#pragma unroll
        for (int i = 0; i < dev::kWaves; ++i) merge(b, sm[3 * i], sm[3 * i + 1], sm[3 * i + 2]);
        part[3 * blockIdx.x] = b.m;
        part[3 * blockIdx.x + 1] = b.s;
        part[3 * blockIdx.x + 2] = b.bad;
    }
    __syncthreads();
}

// Statistics triples of NA partial arrays (np block triples each), computed
// by every wavefront on its own with all loads of a chunk issued together:
// per lane the triples q = lane, lane + 64, ... are merged in order, then
// the wavefront combines.  Every lane receives out[a][0..2].
template <int NA>
__device__ void combine_stats(const double* const (&parts)[NA], int np, double (&out)[NA][3]) {
    const int lane = threadIdx.x % kWave;
    Stat st[NA];
    for (int q0 = 0; q0 < np; q0 += 4 * kWave) {
        double v[NA][4][3];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + lane + u * kWave;
                const bool in = q < np;
                v[a][u][0] = in ? parts[a][3 * q] : 0.0;
                v[a][u][1] = in ? parts[a][3 * q + 1] : 0.0;
                v[a][u][2] = in ? parts[a][3 * q + 2] : 0.0;
            }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int u = 0; u < 4; ++u) merge(st[a], v[a][u][0], v[a][u][1], v[a][u][2]);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        wave_reduce(st[a]);
        out[a][0] = st[a].m;
        out[a][1] = st[a].s;
        out[a][2] = st[a].bad;
    }
}

__device__ __forceinline__ bool needs_second_pass(const double* sb, const double* sa, double factor) {
    const double nb = dev::nominal_norm(sb), na = dev::nominal_norm(sa);
    return isfinite(nb) && isfinite(na) && na < factor * nb;
}

// cs[o] = sum over blocks b of Q[b][o] (o < k) into shared memory, the same
// fixed order in every block and every kernel: thread t takes coefficient
// o = t mod kp (kp = power of two >= the chunk width) and blocks
// b = t / kp, t / kp + S, ... (S = 256 / kp) in order, then a shared-memory
// tree over the S threads of each coefficient.  Loads are issued sixteen at
// a time.  All threads must call it; red holds kBlock doubles.
__device__ void sum_partials(const double* Q, int nbq, int k, double* cs, double* red) {
    const int t = threadIdx.x;
    for (int o0 = 0; o0 < k; o0 += kBlock) {
        const int kc = std::min(kBlock, k - o0);
        int kp = 1;
        while (kp < kc) kp <<= 1;
        const int S = kBlock / kp;
        const int o = t % kp, sidx = t / kp;
        double acc = 0.0;
        if (o < kc) {
            int b = sidx;
            for (; b + 15 * S < nbq; b += 16 * S) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = Q[static_cast<int64_t>(b + u * S) * k + o0 + o];
#pragma unroll
                for (int u = 0; u < 16; ++u) acc += v[u];
            }
            for (; b < nbq; b += S) acc += Q[static_cast<int64_t>(b) * k + o0 + o];
        }
        const double v = dev::reduce_shared_outputs(acc, kp, red);
        if (t < kc) cs[o0 + t] = v;
        __syncthreads();
    }
}

// Per-block projection partials Q[b][o] = sum over the block's rows of
// B[i, o] * y[i], y read from shared memory (ys[row - r0]).
// As project_block for k <= 16 J, with the rows spread over many more
// threads: lanes in groups of 16, lane s of a group accumulates coefficients
// s + 16 j over the rows of its slot (one slot per group, 16 per block),
// four rows per trip; the groups' sums are added by a butterfly within each
// wavefront, then across the four wavefronts in order.
template <typename T, int J>
__device__ void project_block_grouped(const T* __restrict__ B, int64_t ldb, int k, int64_t r0, int nr,
                                      const double* ys, double* red, double* Q) {
    constexpr int LPR = 16, G = kWave / LPR, kSlots = dev::kWaves * G;
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    const int sub = lane % LPR, grp = lane / LPR;
    double acc[J];
#pragma unroll
    for (int j = 0; j < J; ++j) acc[j] = 0.0;
    int i = w * G + grp;
    for (; i + 3 * kSlots < nr; i += 4 * kSlots) {
        double bv[4][J], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int row = i + u * kSlots;
            yv[u] = ys[row];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int o = sub + LPR * j;
                bv[u][j] = o < k ? static_cast<double>(B[(r0 + row) * ldb + o]) : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < J; ++j) acc[j] = fma(bv[u][j], yv[u], acc[j]);
    }
    for (; i < nr; i += kSlots) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int o = sub + LPR * j;
            if (o < k) acc[j] = fma(static_cast<double>(B[(r0 + i) * ldb + o]), ys[i], acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int off = LPR; off < kWave; off <<= 1) acc[j] += __shfl_xor(acc[j], off);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (grp == 0) red[w * LPR + sub] = acc[j];
        __syncthreads();
        if (t < LPR) {
            const int o = t + LPR * j;
            double q = 0.0;
#pragma unroll
            for (int w2 = 0; w2 < dev::kWaves; ++w2) q += red[w2 * LPR + t];
            if (o < k) Q[static_cast<int64_t>(blockIdx.x) * k + o] = q;
        }
        __syncthreads();
    }
}

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
        const double v = dev::reduce_shared_outputs(acc, kp, red);
        if (t < kc) Q[static_cast<int64_t>(blockIdx.x) * k + o0 + t] = v;
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
    // a thread's rows in trips of kTrip: every load of a trip is issued before
    // its stores (r may be written in place), so a trip costs one round trip
    for (int i0 = threadIdx.x; i0 < nr; i0 += kTrip * kBlock) {
        double rv[kTrip], pv[kTrip];
#pragma unroll
        for (int u = 0; u < kTrip; ++u) {
            const int i = i0 + u * kBlock;
            const int64_t g = r0 + i;
            rv[u] = i < nr ? static_cast<double>(r[g * ldr]) : 0.0;
            pv[u] = i < nr ? static_cast<double>(p[g * ldp]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTrip; ++u) {
            const int i = i0 + u * kBlock;
            if (i >= nr) break;
            const int64_t g = r0 + i;
            const T yv = static_cast<T>(rv[u] - pv[u] * a);
            r[g * ldr] = yv;
            ys[i] = static_cast<double>(yv);
            sr.add(rv[u]);
            sy.add(static_cast<double>(yv));
        }
    }
    block_publish(sr, part_r, red);
    block_publish(sy, part_y, red);
    // (the statistics above leave red free after their barrier).  Up to 16
    // columns project_block already gives every thread its own row stream.
    if (k > 16 && k <= 64) project_block_grouped<T, 4>(B, ldb, k, r0, nr, ys, red, Q);
    else if (k > 64 && k <= 128) project_block_grouped<T, 8>(B, ldb, k, r0, nr, ys, red, Q);
    else project_block<T>(B, ldb, k, r0, nr, ys, red, Q);
}

// y -= B c with c = sum_b Qin[b] (gated: second pass only when the DGKS
// test on (before, after) holds); stats of the new y; with Qout, projection
// partials of the new y for a following pass, accumulated in the same read
// of B.  Narrow bases (k <= 16): a thread per row.  Wider bases: a
// wavefront per row, lanes over the k coefficients (J per lane).
template <typename T, int J, int LPR>
__global__ __launch_bounds__(kBlock) void update_kernel(int64_t N, int k, int R, const T* __restrict__ B,
                                                        int64_t ldb, const double* Qin, int nbq, T* y, int64_t ldy,
                                                        double* part_y, double* Qout, const double* go,
                                                        const double* before, const double* after, int np_stats,
                                                        double factor) {
    extern __shared__ double cs[];  // c (k values)
    __shared__ double red[kBlock];
    __shared__ double wq[dev::kWaves][kWave];
    if (*go == 0.0) return;
    if (before != nullptr) {
        const double* const parts[2] = {before, after};
        double st2[2][3];
        combine_stats<2>(parts, np_stats, st2);
        if (!needs_second_pass(st2[0], st2[1], factor)) return;
    }
    sum_partials(Qin, nbq, k, cs, red);
    const int t = threadIdx.x;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    Stat st;
    if (J == 0 && k == 1) {  // one basis column: trips of kTrip rows per thread
        const double c0 = cs[0];
        double q0 = 0.0;
        for (int r1 = t; r1 < nr; r1 += kTrip * kBlock) {
            double bv[kTrip], yo[kTrip];
#pragma unroll
            for (int u = 0; u < kTrip; ++u) {
                const int rr = r1 + u * kBlock;
                const int64_t i = r0 + rr;
                bv[u] = rr < nr ? static_cast<double>(B[i * ldb]) : 0.0;
                yo[u] = rr < nr ? static_cast<double>(y[i * ldy]) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kTrip; ++u) {
                const int rr = r1 + u * kBlock;
                if (rr >= nr) break;
                const T nv = static_cast<T>(yo[u] - fma(bv[u], c0, 0.0));
                y[(r0 + rr) * ldy] = nv;
                st.add(static_cast<double>(nv));
                q0 = fma(bv[u], static_cast<double>(nv), q0);
            }
        }
        block_publish(st, part_y, red);
        if (Qout != nullptr) {
            const int lane = t % kWave, w = t / kWave;
            const double q = dev::wave_sum(q0);
            if (lane == 0) wq[w][0] = q;
            __syncthreads();
            if (t == 0) {
                double qs = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < dev::kWaves; ++w2) qs += wq[w2][0];
                Qout[static_cast<int64_t>(blockIdx.x)] = qs;
            }
        }
        return;
    }
    if (J == 0) {  // k <= 16: a thread per row
        double qa[16];
#pragma unroll
        for (int o = 0; o < 16; ++o) qa[o] = 0.0;
        for (int rr = t; rr < nr; rr += kBlock) {
            const int64_t i = r0 + rr;
            double bv[16];
#pragma unroll
            for (int o = 0; o < 16; ++o) bv[o] = o < k ? static_cast<double>(B[i * ldb + o]) : 0.0;
            double acc = 0.0;
#pragma unroll
            for (int o = 0; o < 16; ++o)
                if (o < k) acc = fma(bv[o], cs[o], acc);
            const T nv = static_cast<T>(static_cast<double>(y[i * ldy]) - acc);
            y[i * ldy] = nv;
            st.add(static_cast<double>(nv));
#pragma unroll
            for (int o = 0; o < 16; ++o) qa[o] = fma(bv[o], static_cast<double>(nv), qa[o]);
        }
        block_publish(st, part_y, red);
        if (Qout != nullptr) {
            // per coefficient: wavefront butterfly, then the kWaves sums in order
            const int lane = t % kWave, w = t / kWave;
            for (int o = 0; o < k; ++o) {
                double q = 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (u == o) q = qa[u];
                q = dev::wave_sum(q);
                if (lane == 0) wq[w][o] = q;
            }
            __syncthreads();
            if (t < k) {
                double q = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < dev::kWaves; ++w2) q += wq[w2][t];
                Qout[static_cast<int64_t>(blockIdx.x) * k + t] = q;
            }
        }
        return;
    }
    // J >= 1: a group of LPR lanes per row (LPR = 64: the whole wavefront),
    // lane s of a group holds coefficients s + LPR j (j < J); a wavefront
    // works on kWave / LPR rows at once, U rows per group per trip
    constexpr int G = kWave / LPR;
    const int lane = t % kWave, w = t / kWave;
    const int sub = lane % LPR, grp = lane / LPR;
    double qa[J > 0 ? J : 1];
    double cl[J > 0 ? J : 1];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        qa[j] = 0.0;
        const int o = sub + LPR * j;
        cl[j] = o < k ? cs[o] : 0.0;
    }
    // U rows per trip, their loads issued together (registers bound U)
    constexpr int U = (J > 0 && (LPR < kWave ? J <= 8 : J <= 4)) ? 4 : 1;
    constexpr int kSlots = dev::kWaves * G;  // rows a block covers per step
    const int slot = w * G + grp;
    int rr = slot;
    for (; rr + (U - 1) * kSlots < nr; rr += U * kSlots) {
        double bv[U][J > 0 ? J : 1], yo[U], acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = r0 + rr + u * kSlots;
            acc[u] = 0.0;
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int o = sub + LPR * j;
                bv[u][j] = o < k ? static_cast<double>(B[i * ldb + o]) : 0.0;
            }
            yo[u] = static_cast<double>(y[i * ldy]);  // every lane of the group reads the same word
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int j = 0; j < J; ++j) acc[u] = fma(bv[u][j], cl[j], acc[u]);
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) acc[u] += __shfl_xor(acc[u], off);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = r0 + rr + u * kSlots;
            const T nv = static_cast<T>(yo[u] - acc[u]);
            if (sub == 0) {
                y[i * ldy] = nv;
                st.add(static_cast<double>(nv));
            }
#pragma unroll
            for (int j = 0; j < J; ++j) qa[j] = fma(bv[u][j], static_cast<double>(nv), qa[j]);
        }
    }
    for (; rr < nr; rr += kSlots) {
        const int64_t i = r0 + rr;
        double bv[J > 0 ? J : 1];
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int o = sub + LPR * j;
            bv[j] = o < k ? static_cast<double>(B[i * ldb + o]) : 0.0;
            acc = fma(bv[j], cl[j], acc);
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        const double yo = static_cast<double>(y[i * ldy]);
        const T nv = static_cast<T>(yo - acc);
        if (sub == 0) {
            y[i * ldy] = nv;
            st.add(static_cast<double>(nv));
        }
#pragma unroll
        for (int j = 0; j < J; ++j) qa[j] = fma(bv[j], static_cast<double>(nv), qa[j]);
    }
    block_publish(st, part_y, red);
    if (Qout != nullptr) {
        // the G groups of a wavefront add their sums (butterfly over the group
        // bits), then Qout[block][o] = the kWaves wavefront sums in order
#pragma unroll
        for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int off = LPR; off < kWave; off <<= 1) qa[j] += __shfl_xor(qa[j], off);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (grp == 0) wq[w][sub] = qa[j];
            __syncthreads();
            if (t < LPR) {
                const int o = t + LPR * j;
                double q = 0.0;
#pragma unroll
                for (int w2 = 0; w2 < dev::kWaves; ++w2) q += wq[w2][t];
                if (o < k) Qout[static_cast<int64_t>(blockIdx.x) * k + o] = q;
            }
            __syncthreads();
        }
    }
}

// Rank decision and normalization of y (the host rule for a one-column
// block): ||scale|| from part_scale, ||resid|| from the first-pass partials,
// or from the second-pass ones when the DGKS test on (before, first) held.
// Accepted when both are finite and ||resid|| > rank_factor * ||scale||;
// then dst = y * (1 / ||resid||) and, with b given, gpart[block] = sum over
// the block of dst[i] b[i].  Block 0 writes record = {resid, scale,
// accepted, c} (c = summed coefficient of a one-column basis, when Q1 is
// given, combined exactly as the update kernels combined it) and, on
// acceptance, *cell = resid; on rejection *go = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void chain_normalize_kernel(
    int64_t N, int R, const T* y, int64_t ldy, T* dst, int64_t ldd, const double* part_scale,
    const double* part_before, const double* part_first, const double* part_second, int np_stats,
    double dgks_factor, double rank_factor, const double* Q1, const double* Q2, int nbq, const double* b,
    double* gpart, double* record, double* cell, double* go) {
    __shared__ double red[kBlock];
    __shared__ double cq[2];
    if (*go == 0.0) return;
    const double* const parts[4] = {part_scale, part_before, part_first, part_second};
    double st4[4][3];
    combine_stats<4>(parts, np_stats, st4);
    const bool second = needs_second_pass(st4[1], st4[2], dgks_factor);
    const double scale = dev::nominal_norm(st4[0]), resid = dev::nominal_norm(second ? st4[3] : st4[2]);
    const bool ok = isfinite(scale) && scale >= 0.0 && isfinite(resid) && resid > rank_factor * scale;
    if (blockIdx.x == 0) {
        if (Q1 != nullptr) {
            sum_partials(Q1, nbq, 1, cq, red);
            if (second) sum_partials(Q2, nbq, 1, cq + 1, red);
        }
        if (threadIdx.x == 0) {
            record[0] = resid;
            record[1] = scale;
            record[2] = ok ? 1.0 : 0.0;
            // the coefficient sums of a one-column basis, as the updates used them
            if (Q1 != nullptr) record[3] = second ? cq[0] + cq[1] : cq[0];
            if (ok) *cell = resid;
            else *go = 0.0;
        }
    }
    if (!ok) return;
    const double inv = 1.0 / resid;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int nr = static_cast<int>(std::min<int64_t>(R, N - r0));
    double acc = 0.0;
    // trips of kTrip rows: loads before stores (dst may be y itself)
    for (int i0 = threadIdx.x; i0 < nr; i0 += kTrip * kBlock) {
        double v[kTrip], bv[kTrip];
#pragma unroll
        for (int u = 0; u < kTrip; ++u) {
            const int i = i0 + u * kBlock;
            const int64_t g = r0 + i;
            v[u] = i < nr ? static_cast<double>(y[g * ldy]) : 0.0;
            bv[u] = (i < nr && b != nullptr) ? b[g] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTrip; ++u) {
            const int i = i0 + u * kBlock;
            if (i >= nr) break;
            const T nv = static_cast<T>(inv == 1.0 ? v[u] : inv * v[u]);
            dst[(r0 + i) * ldd] = nv;
            if (b != nullptr) acc = fma(static_cast<double>(nv), bv[u], acc);
        }
    }
    if (b != nullptr) {
        acc = dev::block_sum(acc, red);
        if (threadIdx.x == 0) gpart[blockIdx.x] = acc;
    }
}

// Q[0][o] = sum over the nb blocks of Q[b][o], in place: thread (o, g) adds
// blocks g, g + 8, ... in order, then the eight group sums are added in group
// order; a block reads and writes only its own 32 columns (row 0 of a column
// is read by its group 0 before the barrier and written after it).
constexpr int kCombineCols = 32, kCombineGroups = kBlock / kCombineCols;
__global__ __launch_bounds__(kBlock) void chain_combine_kernel(double* Q, int nb, int k, const double* go) {
    if (*go == 0.0) return;
    __shared__ double red[kCombineGroups][kCombineCols];
    const int ol = threadIdx.x % kCombineCols, g = threadIdx.x / kCombineCols;
    const int o = blockIdx.x * kCombineCols + ol;
    double acc = 0.0;
    if (o < k) {
        int b = g;
        for (; b + 3 * kCombineGroups < nb; b += 4 * kCombineGroups) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = Q[static_cast<int64_t>(b + u * kCombineGroups) * k + o];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u];
        }
        for (; b < nb; b += kCombineGroups) acc += Q[static_cast<int64_t>(b) * k + o];
    }
    red[g][ol] = acc;
    __syncthreads();
    if (g == 0 && o < k) {
        double v = red[0][ol];
#pragma unroll
        for (int q = 1; q < kCombineGroups; ++q) v += red[q][ol];
        Q[o] = v;
    }
}

}  // namespace

// Whether the nb x k projection partials of the chain kernels on N rows pass
// through chain_combine_kernel before their consumer (wide bases: k > 256)
// instead of every consumer block re-reading them all.
static bool chain_combines(int64_t /*N*/, int k) { return k > 256; }

int chain_block_rows(int64_t N, int k) {
    // about 128 blocks; for k <= 256, fewer and longer blocks when the k
    // projection partials each consumer block combines would exceed 8K values
    const int64_t base = std::max<int64_t>(64, ceil_div(N, 128));
    const int64_t wide = chain_combines(N, k) ? base : ceil_div(N * std::max(k, 1), 8192);
    return static_cast<int>(std::min<int64_t>(4096, std::max(base, wide)));
}

static void chain_combine(hipStream_t s, double* Q, int64_t nb, int k, const double* go) {
    hipLaunchKernelGGL(chain_combine_kernel, dim3(static_cast<unsigned>(ceil_div(k, kCombineCols))), dim3(kBlock), 0, s,
                       Q, static_cast<int>(nb), k, go);
    AMFLS_LAUNCH_CHECK();
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
    if (chain_combines(N, k)) chain_combine(s, Q, nb, k, go);
}

template <typename T>
void chain_update(hipStream_t s, int64_t N, int k, const T* B, int64_t ldb, const double* Qin, T* y, int64_t ldy,
                  double* part_y, double* Qout, const double* go, const double* before, const double* after,
                  double factor) {
    if (k < 1 || k > 2048) throw std::invalid_argument("chain update supports bases of 1 to 2048 columns");
    const int R = chain_block_rows(N, k);
    const int64_t nb = ceil_div(N, R);
    const size_t lds = static_cast<size_t>(k) * sizeof(double);
    const dim3 grid(static_cast<unsigned>(nb));
    const bool combined = chain_combines(N, k);  // Qin holds its sums in row 0
    const int nbq = combined ? 1 : static_cast<int>(nb);
#define AMFLS_UPDATE(J, LPR)                                                                                     \
    hipLaunchKernelGGL((update_kernel<T, J, LPR>), grid, dim3(kBlock), lds, s, N, k, R, B, ldb, Qin, nbq, y, ldy, \
                       part_y, Qout, go, before, after, static_cast<int>(nb), factor)
    // up to 128 columns: groups of 16 lanes per row (four rows per wavefront
    // at once); wider bases: a wavefront per row
    if (k <= 16) AMFLS_UPDATE(0, 64);
    else if (k <= 64) AMFLS_UPDATE(4, 16);
    else if (k <= 128) AMFLS_UPDATE(8, 16);
    else if (k <= 256) AMFLS_UPDATE(4, 64);
    else if (k <= 512) AMFLS_UPDATE(8, 64);
    else if (k <= 1024) AMFLS_UPDATE(16, 64);
    else AMFLS_UPDATE(32, 64);
#undef AMFLS_UPDATE
    AMFLS_LAUNCH_CHECK();
    if (Qout != nullptr && combined) chain_combine(s, Qout, nb, k, go);
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
                       static_cast<int>(nb), dgks_factor,
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
