// CSR block products for the AMFLS engine on gfx950.
//
// Y = A X with X, Y row-major (n x s and m x s).  A group of G = P * LC
// lanes owns one CSR row: LC lanes split the s block columns (each lane
// holds CPL columns) and P lanes split the row's entries; the P partial sums
// are combined with xor-shuffles inside the (wave-aligned) group.  A row-major
// block makes every nonzero a contiguous s-wide gather from X, so one
// nonzero costs one coalesced row segment instead of s scattered words.
//
// A^T Y uses the same kernel on a transposed CSR that is built once on the
// device by a stable radix sort of the column indices, so no atomics are
// needed and the summation order is deterministic.
#include "kernels.hpp"

#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace amflsx {
namespace {

constexpr int kThreads = 256;

// In-kernel product timing (ProductStamp): block 0 of the first launch
// stamps the start; every block of the final launch, once all its threads
// are done, raises its end slot (one of kStampEnds, by block index) to its
// own end time with an atomic maximum, so the largest end slot is the last
// block to finish whatever the row lengths (the clock only grows, so a value
// left by an earlier product is smaller).
__device__ __forceinline__ void stamp_begin(ProductStamp st) {
    if (st.slots != nullptr && st.first && blockIdx.x == 0 && threadIdx.x == 0) st.slots[0] = wall_clock64();
}

__device__ __forceinline__ void stamp_end(ProductStamp st) {
    if (st.slots == nullptr || !st.last) return;  // uniform: the barrier sees every thread or none
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(st.slots + kStampStride * (1 + static_cast<int64_t>(blockIdx.x % kStampEnds)),
                  static_cast<unsigned long long>(wall_clock64()));
}

template <typename V, typename T, int P, int LC, int CPL>
__device__ __forceinline__ void spmm_rows(int64_t rows, const int64_t* __restrict__ rp,
                                          const int32_t* __restrict__ ci, const V* __restrict__ val, int ncols,
                                          const T* __restrict__ X, int64_t ldx, T* __restrict__ Y, int64_t ldy,
                                          const T* __restrict__ Add, int64_t ld_add, T add_scale) {
    constexpr int G = P * LC;
    static_assert(G <= kWave && (kWave % G) == 0, "group must tile a wavefront");
    const int gid = threadIdx.x / G;
    const int lig = threadIdx.x % G;
    const int sub = lig / LC;
    const int cl = lig % LC;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * (kThreads / G) + gid;
    if (row >= rows) return;  // whole group leaves together
    T acc[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) acc[t] = T(0);
    const int64_t e0 = rp[row], e1 = rp[row + 1];
    int64_t e = e0 + sub;
    // four entries per trip: their indices and values, then their gathers,
    // are independent loads in flight together (entries still summed in order)
    for (; e + 3 * P < e1; e += 4 * P) {
        int32_t c[4];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = ci[e + u * P];
            v[u] = load_as<T, V>(val[e + u * P]);
        }
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int j = cl + LC * t;
            if (j < ncols) {
                T xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) xv[u] = X[static_cast<int64_t>(c[u]) * ldx + j];
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[t] = fma(v[u], xv[u], acc[t]);
            }
        }
    }
    for (; e + P < e1; e += 2 * P) {
        const int32_t c0 = ci[e], c1 = ci[e + P];
        const T v0 = load_as<T, V>(val[e]);
        const T v1 = load_as<T, V>(val[e + P]);
        const T* x0 = X + static_cast<int64_t>(c0) * ldx;
        const T* x1 = X + static_cast<int64_t>(c1) * ldx;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int j = cl + LC * t;
            if (j < ncols) {
                acc[t] = fma(v0, x0[j], acc[t]);
                acc[t] = fma(v1, x1[j], acc[t]);
            }
        }
    }
    if (e < e1) {
        const int32_t c0 = ci[e];
        const T v0 = load_as<T, V>(val[e]);
        const T* x0 = X + static_cast<int64_t>(c0) * ldx;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int j = cl + LC * t;
            if (j < ncols) acc[t] = fma(v0, x0[j], acc[t]);
        }
    }
    if constexpr (P > 1) {
#pragma unroll
        for (int off = LC; off < G; off <<= 1) {
#pragma unroll
            for (int t = 0; t < CPL; ++t) acc[t] += __shfl_xor(acc[t], off);
        }
    }
    if (sub != 0) return;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = cl + LC * t;
        if (j >= ncols) continue;
        T r = acc[t];
        if (Add != nullptr) r = fma(add_scale, Add[row * ld_add + j], r);
        Y[row * ldy + j] = r;
    }
}

template <typename V, typename T, int P, int LC, int CPL>
__global__ __launch_bounds__(kThreads) void csr_spmm_kernel(
    int64_t rows, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const V* __restrict__ val, int ncols, const T* __restrict__ X, int64_t ldx, T* __restrict__ Y,
    int64_t ldy, const T* __restrict__ Add, int64_t ld_add, T add_scale, ProductStamp stamp) {
    stamp_begin(stamp);
    spmm_rows<V, T, P, LC, CPL>(rows, rp, ci, val, ncols, X, ldx, Y, ldy, Add, ld_add, add_scale);
    stamp_end(stamp);
}

// Width-one products by CSR-Stream (J. L. Greathouse and M. Daga, "Efficient
// sparse matrix-vector multiplication on GPUs using the CSR storage format",
// SC14): a block owns consecutive rows holding at most B entries; its
// threads stream those entries with coalesced loads, multiply by the
// gathered x and park the products in LDS; then a thread per row sums its
// products in entry order.  A row longer than the budget gets a block that
// reduces it with all threads.  Two budgets: 2048 entries (8 per thread) and
// 8192 (32 per thread, 4096 rows at most).  A block's life is a fixed chain
// of dependent memory trips (partition, row pointers, entries, gathers,
// sums), so on a matrix of millions of rows the wide budget does four times
// the work per chain with a quarter of the blocks.
constexpr int kStreamBudget = 2048;
constexpr int kStreamMaxRows = 1024;
constexpr int kStreamWideBudget = 8192;
constexpr int kStreamWideMaxRows = 4096;

template <typename V, typename T, int B>
__device__ __forceinline__ void stream_block(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                             const V* __restrict__ val, const int64_t* __restrict__ br,
                                             const T* __restrict__ X, int64_t ldx, T* __restrict__ Y, int64_t ldy,
                                             const T* __restrict__ Add, int64_t ld_add, T add_scale, T* prod,
                                             T* red) {
    const int t = threadIdx.x;
    const int64_t r0 = br[blockIdx.x], r1 = br[blockIdx.x + 1];
    const int64_t e0 = rp[r0], e1 = rp[r1];
    if (e1 - e0 <= B) {
        // all of a thread's entries in two waves of independent loads: the
        // column indices and values, then the gathered x
        constexpr int U = B / kThreads;
        int32_t c[U];
        T v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = e0 + t + u * kThreads;
            const bool in = e < e1;
            c[u] = in ? ci[e] : 0;
            v[u] = in ? load_as<T, V>(val[e]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = e0 + t + u * kThreads;
            xv[u] = e < e1 ? X[static_cast<int64_t>(c[u]) * ldx] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = e0 + t + u * kThreads;
            if (e < e1) prod[e - e0] = v[u] * xv[u];
        }
        __syncthreads();
        for (int64_t r = r0 + t; r < r1; r += kThreads) {
            const int a = static_cast<int>(rp[r] - e0), b = static_cast<int>(rp[r + 1] - e0);
            T acc = T(0);
            for (int i = a; i < b; ++i) acc += prod[i];
            if (Add != nullptr) acc = fma(add_scale, Add[r * ld_add], acc);
            Y[r * ldy] = acc;
        }
        return;
    }
    // one row longer than the budget (r1 == r0 + 1)
    T acc = T(0);
    for (int64_t e = e0 + t; e < e1; e += kThreads)
        acc = fma(load_as<T, V>(val[e]), X[static_cast<int64_t>(ci[e]) * ldx], acc);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (t % kWave == 0) red[t / kWave] = acc;
    __syncthreads();
    if (t == 0) {
        T tot = T(0);
#pragma unroll
        for (int w = 0; w < kThreads / kWave; ++w) tot += red[w];
        if (Add != nullptr) tot = fma(add_scale, Add[r0 * ld_add], tot);
        Y[r0 * ldy] = tot;
    }
}

template <typename V, typename T, int B>
__global__ __launch_bounds__(kThreads) void csr_stream_kernel(
    const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, const V* __restrict__ val,
    const int64_t* __restrict__ br, const T* __restrict__ X, int64_t ldx, T* __restrict__ Y, int64_t ldy,
    const T* __restrict__ Add, int64_t ld_add, T add_scale, ProductStamp stamp) {
    __shared__ T prod[B];
    __shared__ T red[kThreads / kWave];
    stamp_begin(stamp);
    stream_block<V, T, B>(rp, ci, val, br, X, ldx, Y, ldy, Add, ld_add, add_scale, prod, red);
    stamp_end(stamp);
}

// Width-one products with x staged in LDS, on the slab-tiled copy (a
// panel's entries grouped by column slab of kSlabCols columns, then by row
// and column).  A block owns a panel.  For each slab that holds some of the
// panel's entries, it stages the slab's x in LDS with coalesced loads; each
// entry then reads its x from LDS (a random 8-byte LDS read instead of a
// cache line through the vector L1, which bounds the row kernels' gathers at
// about one per clock per CU), and a run of one row's products is summed in
// entry order and added to the row's sum in LDS.  The next slab's x and its
// first kSlabAhead chunks of entries are loaded into registers while the
// current slab is processed.  A row's sum is ((run in its first slab) + run
// in the next slab) + ..., each run summed in column order (a run split by a
// chunk boundary adds its two parts in order): deterministic, but rounded
// differently from the row kernels' single fma chain.
constexpr int kSlabCols = 8192;       // columns per slab
constexpr int kSlabRows = 8192;       // rows per panel at most (16-bit row offsets)
constexpr int kSlabMaxCount = 1024;   // slabs per matrix at most
constexpr int kSlabThreads = 1024;
constexpr int kSlabU = 2;             // entries per thread per chunk
constexpr int kSlabChunk = kSlabThreads * kSlabU;
constexpr int kSlabAhead = 2;         // chunks of the next slab loaded ahead
constexpr int kSlabStage = kSlabCols / kSlabThreads;
constexpr int kSlabPanels = 256;      // panels per matrix, about one per CU
static_assert(kSlabCols <= 65536 && kSlabRows <= 65536, "packed 16-bit offsets");

// Barrier for LDS traffic only: waits for this wave's LDS operations and
// never for its global loads, so the next slab's loads stay in flight
// across it whatever fences the compiler would attach to __syncthreads().
__device__ __forceinline__ void slab_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <typename T>
struct SlabEntries {
    uint32_t i[kSlabAhead][kSlabU];
    T v[kSlabAhead][kSlabU];
};

// Loads of slab k: its x (kSlabStage values per thread, zero past the last
// column) and the first kSlabAhead chunks of the tile's n entries at `base`.
template <typename V, typename T>
__device__ __forceinline__ void slab_fetch(int k, int64_t cols, const T* __restrict__ X, int64_t ldx,
                                           const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                           int64_t base, int n, T (&x)[kSlabStage], SlabEntries<T>& e) {
    const int t = threadIdx.x;
    const int64_t c0 = static_cast<int64_t>(k) * kSlabCols;
#pragma unroll
    for (int u = 0; u < kSlabStage; ++u) {
        const int64_t c = c0 + t + u * kSlabThreads;
        x[u] = c < cols ? X[c * ldx] : T(0);
    }
#pragma unroll
    for (int q = 0; q < kSlabAhead; ++q) {
#pragma unroll
        for (int u = 0; u < kSlabU; ++u) {
            const int pos = q * kSlabChunk + t + u * kSlabThreads;
            const bool in = pos < n;
            e.i[q][u] = in ? idx[base + pos] : 0u;
            e.v[q][u] = in ? load_as<T, V>(val[base + pos]) : T(0);
        }
    }
}

// Cross-lane moves (DPP) of one 32-bit value; lanes whose source lies
// outside the pattern, or outside the rows that ROW_MASK selects, get `old`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t x) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(x), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_zero(float x) {
    return __uint_as_float(dpp_u32<CTRL, ROW_MASK>(0u, __float_as_uint(x)));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_zero(double x) {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(x));
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>(0u, static_cast<uint32_t>(b));
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>(0u, static_cast<uint32_t>(b >> 32));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// One step of the segmented scan: add the partial sum of the lane the
// pattern names when that lane holds the same key (keys are sorted, so every
// lane between the two holds it too).
template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ void seg_step(T& v, uint32_t key) {
    const uint32_t k = dpp_u32<CTRL, ROW_MASK>(0xFFFFFFFFu, key);
    const T u = dpp_zero<CTRL, ROW_MASK>(v);
    if (k == key) v += u;
}

// Segmented inclusive sum over a wavefront of sorted keys: afterwards lane
// i holds the sum of its key's values from the key's first lane in the
// wavefront to lane i.  Shifts by 1, 2, 4 and 8 lanes inside each row of 16
// lanes, then lane 15 of each row into the next row (rows 1 and 3) and lane
// 31 into rows 2 and 3 (gfx9 row broadcasts): a fixed tree, so the sum's
// rounding is the same on every run.
template <typename T>
__device__ __forceinline__ void wave_segmented_scan(T& v, uint32_t key) {
    seg_step<0x111, 0xF>(v, key);  // row_shr:1
    seg_step<0x112, 0xF>(v, key);  // row_shr:2
    seg_step<0x114, 0xF>(v, key);  // row_shr:4
    seg_step<0x118, 0xF>(v, key);  // row_shr:8
    seg_step<0x142, 0xA>(v, key);  // row_bcast:15
    seg_step<0x143, 0xC>(v, key);  // row_bcast:31
}

// One chunk of n <= kSlabChunk entries of the current tile.  Entry `pos`
// sits in group g = pos / 64 (one wavefront's slot: g = u * waves + wave)
// at lane pos % 64.  Each group sums its rows' runs with a segmented scan;
// the lane ending a run owns it unless the run began in an earlier group,
// adds the head runs of the following groups that continue it (published in
// LDS), and adds the total to the row's sum: one owner per row and chunk,
// and a row's parts are added in entry order.  Ends with a barrier.
constexpr int kSlabGroups = kSlabChunk / kWave;
constexpr uint32_t kSlabNoKey = 0xFFFFFFFEu;  // keys of positions past the chunk

template <typename T>
__device__ __forceinline__ void slab_chunk(int n, const uint32_t (&ii)[kSlabU], const T (&vv)[kSlabU],
                                           const T* xs, T* acc, T* hsum, uint32_t* hkey, int* hend,
                                           uint32_t* tkey) {
    const int lane = threadIdx.x % kWave;
    const int w = threadIdx.x / kWave;
    T v[kSlabU];
    uint32_t key[kSlabU];
    bool end[kSlabU];
    int e0[kSlabU];
#pragma unroll
    for (int u = 0; u < kSlabU; ++u) {
        const int pos = (u * (kSlabThreads / kWave) + w) * kWave + lane;
        const bool in = pos < n;
        v[u] = in ? vv[u] * xs[ii[u] & 0xFFFFu] : T(0);
        key[u] = in ? (ii[u] >> 16) : kSlabNoKey;
    }
#pragma unroll
    for (int u = 0; u < kSlabU; ++u) wave_segmented_scan(v[u], key[u]);
#pragma unroll
    for (int u = 0; u < kSlabU; ++u) {
        const int g = u * (kSlabThreads / kWave) + w;
        const uint32_t next = dpp_u32<0x130, 0xF>(0xFFFFFFFFu, key[u]);  // wave_shl:1, lane i + 1's key
        end[u] = key[u] != kSlabNoKey && next != key[u];
        const unsigned long long ends = __ballot(end[u]);
        e0[u] = ends != 0ull ? __ffsll(ends) - 1 : kWave;
        if (end[u] && lane == e0[u]) {
            hsum[g] = v[u];
            hkey[g] = key[u];
            hend[g] = e0[u];
        }
        if (key[u] != kSlabNoKey && (lane == kWave - 1 || next == kSlabNoKey)) tkey[g] = key[u];
    }
    slab_barrier();
#pragma unroll
    for (int u = 0; u < kSlabU; ++u) {
        if (!end[u]) continue;
        const int g = u * (kSlabThreads / kWave) + w;
        // a head run that continues the previous group's last run is added by that run's owner
        if (lane == e0[u] && g > 0 && tkey[g - 1] == key[u]) continue;
        T total = v[u];
        if (lane == kWave - 1) {
            for (int g2 = g + 1; g2 < kSlabGroups && g2 * kWave < n && hkey[g2] == key[u]; ++g2) {
                total += hsum[g2];
                if (hend[g2] != kWave - 1) break;
            }
        }
        acc[key[u]] += total;
    }
    slab_barrier();
}

template <typename V, typename T>
__global__ __launch_bounds__(kSlabThreads) void csr_slab_kernel(
    int64_t cols, int nslab, const int64_t* __restrict__ panels, const int64_t* __restrict__ tiles,
    const uint32_t* __restrict__ idx, const V* __restrict__ val, const T* __restrict__ X, int64_t ldx,
    T* __restrict__ Y, int64_t ldy, const T* __restrict__ Add, int64_t ld_add, T add_scale, ProductStamp stamp) {
    __shared__ T xs[kSlabCols];
    __shared__ T acc[kSlabRows];
    __shared__ T hsum[kSlabGroups];
    __shared__ uint32_t hkey[kSlabGroups];
    __shared__ int hend[kSlabGroups];
    __shared__ uint32_t tkey[kSlabGroups];
    __shared__ int32_t toff[kSlabMaxCount + 1];
    stamp_begin(stamp);
    const int t = threadIdx.x;
    const int64_t p = blockIdx.x;
    const int64_t r0 = panels[p];
    const int nrows = static_cast<int>(panels[p + 1] - r0);
    const int64_t tb = p * nslab;
    const int64_t pe0 = tiles[tb];
    for (int k = t; k <= nslab; k += kSlabThreads) toff[k] = static_cast<int32_t>(tiles[tb + k] - pe0);
    for (int i = t; i < nrows; i += kSlabThreads) acc[i] = T(0);
    slab_barrier();
    int k = 0;
    while (k < nslab && toff[k + 1] == toff[k]) ++k;
    T xn[kSlabStage];
    SlabEntries<T> en;
    if (k < nslab) slab_fetch<V, T>(k, cols, X, ldx, idx, val, pe0 + toff[k], toff[k + 1] - toff[k], xn, en);
    while (k < nslab) {
        // the previous tile's last chunk ended with a barrier: xs is free
#pragma unroll
        for (int u = 0; u < kSlabStage; ++u) xs[t + u * kSlabThreads] = xn[u];
        const SlabEntries<T> cur = en;
        const int64_t base = pe0 + toff[k];
        const int n = toff[k + 1] - toff[k];
        int kn = k + 1;
        while (kn < nslab && toff[kn + 1] == toff[kn]) ++kn;
        if (kn < nslab)
            slab_fetch<V, T>(kn, cols, X, ldx, idx, val, pe0 + toff[kn], toff[kn + 1] - toff[kn], xn, en);
        slab_barrier();
#pragma unroll
        for (int q = 0; q < kSlabAhead; ++q)
            if (q * kSlabChunk < n)
                slab_chunk<T>(min(kSlabChunk, n - q * kSlabChunk), cur.i[q], cur.v[q], xs, acc, hsum, hkey, hend, tkey);
        for (int c = kSlabAhead * kSlabChunk; c < n; c += kSlabChunk) {
            uint32_t ii[kSlabU];
            T vv[kSlabU];
#pragma unroll
            for (int u = 0; u < kSlabU; ++u) {
                const int pos = c + t + u * kSlabThreads;
                const bool in = pos < n;
                ii[u] = in ? idx[base + pos] : 0u;
                vv[u] = in ? load_as<T, V>(val[base + pos]) : T(0);
            }
            slab_chunk<T>(min(kSlabChunk, n - c), ii, vv, xs, acc, hsum, hkey, hend, tkey);
        }
        k = kn;
    }
    __syncthreads();  // every row's sum is complete (also for a panel without entries)
    for (int i = t; i < nrows; i += kSlabThreads) {
        const int64_t r = r0 + i;
        T out = acc[i];
        if (Add != nullptr) out = fma(add_scale, Add[r * ld_add], out);
        Y[r * ldy] = out;
    }
    stamp_end(stamp);
}

struct VariantInfo {
    int p, lc, cpl;
    const char* name;
};

// (entry lanes P, column lanes LC, columns per lane CPL)
constexpr VariantInfo kVariants[] = {
    {1, 1, 1, "p1_c1"},
    {2, 1, 1, "p2_c1"},
    {4, 1, 1, "p4_c1"},
    {8, 1, 1, "p8_c1"},
    {16, 1, 1, "p16_c1"},
    {32, 1, 1, "p32_c1"},
    {64, 1, 1, "p64_c1"},
    {1, 2, 1, "p1_c2"},
    {2, 2, 1, "p2_c2"},
    {4, 2, 1, "p4_c2"},
    {8, 2, 1, "p8_c2"},
    {16, 2, 1, "p16_c2"},
    {32, 2, 1, "p32_c2"},
    {1, 4, 1, "p1_c4"},
    {2, 4, 1, "p2_c4"},
    {4, 4, 1, "p4_c4"},
    {8, 4, 1, "p8_c4"},
    {16, 4, 1, "p16_c4"},
    {1, 8, 1, "p1_c8"},
    {2, 8, 1, "p2_c8"},
    {4, 8, 1, "p4_c8"},
    {8, 8, 1, "p8_c8"},
    {1, 16, 1, "p1_c16"},
    {2, 16, 1, "p2_c16"},
    {4, 16, 1, "p4_c16"},
    {1, 32, 1, "p1_c32"},
    {2, 32, 1, "p2_c32"},
    {1, 64, 1, "p1_c64"},
    {1, 64, 2, "p1_c128"},
    {1, 64, 4, "p1_c256"},
};
constexpr int kVariantCount = sizeof(kVariants) / sizeof(kVariants[0]);

template <typename V, typename T, int P, int LC, int CPL>
void launch_variant(hipStream_t s, const CsrView& A, int ncols, const T* X, int64_t ldx, T* Y,
                    int64_t ldy, const T* Add, int64_t ld_add, double add_scale, hipEvent_t e0, hipEvent_t e1,
                    ProductStamp st) {
    constexpr int G = P * LC;
    const int64_t groups_per_block = kThreads / G;
    const dim3 grid(static_cast<unsigned>(ceil_div(A.rows, groups_per_block)));
    if (e0 || e1)
        hipExtLaunchKernelGGL((csr_spmm_kernel<V, T, P, LC, CPL>), grid, dim3(kThreads), 0, s, e0, e1, 0, A.rows,
                              A.row_ptr, A.col_idx, static_cast<const V*>(A.values), ncols, X, ldx, Y, ldy, Add,
                              ld_add, static_cast<T>(add_scale), st);
    else
        hipLaunchKernelGGL((csr_spmm_kernel<V, T, P, LC, CPL>), grid, dim3(kThreads), 0, s, A.rows, A.row_ptr,
                           A.col_idx, static_cast<const V*>(A.values), ncols, X, ldx, Y, ldy, Add, ld_add,
                           static_cast<T>(add_scale), st);
    AMFLS_LAUNCH_CHECK();
}

template <typename V, typename T, int B>
void launch_stream_typed(hipStream_t s, const CsrView& A, const T* X, int64_t ldx, T* Y, int64_t ldy,
                         const T* Add, int64_t ld_add, double add_scale, hipEvent_t e0, hipEvent_t e1,
                         ProductStamp st) {
    const bool wide = B == kStreamWideBudget;
    const int64_t* br = wide ? A.stream_rows_wide : A.stream_rows;
    const dim3 grid(static_cast<unsigned>(wide ? A.stream_blocks_wide : A.stream_blocks));
    if (e0 || e1)
        hipExtLaunchKernelGGL((csr_stream_kernel<V, T, B>), grid, dim3(kThreads), 0, s, e0, e1, 0, A.row_ptr,
                              A.col_idx, static_cast<const V*>(A.values), br, X, ldx, Y, ldy, Add, ld_add,
                              static_cast<T>(add_scale), st);
    else
        hipLaunchKernelGGL((csr_stream_kernel<V, T, B>), grid, dim3(kThreads), 0, s, A.row_ptr, A.col_idx,
                           static_cast<const V*>(A.values), br, X, ldx, Y, ldy, Add, ld_add,
                           static_cast<T>(add_scale), st);
    AMFLS_LAUNCH_CHECK();
}

template <typename T, int B>
void launch_stream(hipStream_t s, const CsrView& A, const T* X, int64_t ldx, T* Y, int64_t ldy, const T* Add,
                   int64_t ld_add, double add_scale, hipEvent_t e0, hipEvent_t e1, ProductStamp st) {
    if (A.type == ValueType::f32)
        launch_stream_typed<float, T, B>(s, A, X, ldx, Y, ldy, Add, ld_add, add_scale, e0, e1, st);
    else if (A.type == ValueType::f64)
        launch_stream_typed<double, T, B>(s, A, X, ldx, Y, ldy, Add, ld_add, add_scale, e0, e1, st);
    else
        throw std::invalid_argument("CSR values must be float32 or float64");
}

template <typename V, typename T>
void launch_slab_typed(hipStream_t s, const CsrView& A, const T* X, int64_t ldx, T* Y, int64_t ldy, const T* Add,
                       int64_t ld_add, double add_scale, hipEvent_t e0, hipEvent_t e1, ProductStamp st) {
    const dim3 grid(static_cast<unsigned>(A.slab_panel_count));
    const V* val = static_cast<const V*>(A.slab_val);
    if (e0 || e1)
        hipExtLaunchKernelGGL((csr_slab_kernel<V, T>), grid, dim3(kSlabThreads), 0, s, e0, e1, 0, A.cols,
                              A.slab_count, A.slab_panels, A.slab_tiles, A.slab_idx, val, X, ldx, Y, ldy, Add,
                              ld_add, static_cast<T>(add_scale), st);
    else
        hipLaunchKernelGGL((csr_slab_kernel<V, T>), grid, dim3(kSlabThreads), 0, s, A.cols, A.slab_count,
                           A.slab_panels, A.slab_tiles, A.slab_idx, val, X, ldx, Y, ldy, Add, ld_add,
                           static_cast<T>(add_scale), st);
    AMFLS_LAUNCH_CHECK();
}

template <typename T>
void launch_slab(hipStream_t s, const CsrView& A, const T* X, int64_t ldx, T* Y, int64_t ldy, const T* Add,
                 int64_t ld_add, double add_scale, hipEvent_t e0, hipEvent_t e1, ProductStamp st) {
    if (A.type == ValueType::f32)
        launch_slab_typed<float, T>(s, A, X, ldx, Y, ldy, Add, ld_add, add_scale, e0, e1, st);
    else if (A.type == ValueType::f64)
        launch_slab_typed<double, T>(s, A, X, ldx, Y, ldy, Add, ld_add, add_scale, e0, e1, st);
    else
        throw std::invalid_argument("CSR values must be float32 or float64");
}

template <typename V, typename T>
void launch_by_index(int variant, hipStream_t s, const CsrView& A, int ncols, const T* X, int64_t ldx,
                     T* Y, int64_t ldy, const T* Add, int64_t ld_add, double add_scale, hipEvent_t e0,
                     hipEvent_t e1, ProductStamp st) {
#define AMFLS_CASE(i)                                                                            \
    case i:                                                                                      \
        launch_variant<V, T, kVariants[i].p, kVariants[i].lc, kVariants[i].cpl>(                 \
            s, A, ncols, X, ldx, Y, ldy, Add, ld_add, add_scale, e0, e1, st);              \
        break;
    static_assert(kVariantCount == 30, "variant switch must cover the table");
    switch (variant) {
        AMFLS_CASE(0) AMFLS_CASE(1) AMFLS_CASE(2) AMFLS_CASE(3) AMFLS_CASE(4) AMFLS_CASE(5) AMFLS_CASE(6)
        AMFLS_CASE(7) AMFLS_CASE(8) AMFLS_CASE(9) AMFLS_CASE(10) AMFLS_CASE(11) AMFLS_CASE(12)
        AMFLS_CASE(13) AMFLS_CASE(14) AMFLS_CASE(15) AMFLS_CASE(16) AMFLS_CASE(17) AMFLS_CASE(18)
        AMFLS_CASE(19) AMFLS_CASE(20) AMFLS_CASE(21) AMFLS_CASE(22) AMFLS_CASE(23) AMFLS_CASE(24)
        AMFLS_CASE(25) AMFLS_CASE(26) AMFLS_CASE(27) AMFLS_CASE(28) AMFLS_CASE(29)
        default:
            throw std::invalid_argument("unknown CSR SpMM variant");
    }
#undef AMFLS_CASE
}

// ------------------------------------------------------------ transpose
__global__ void row_of_kernel(int64_t rows, const int64_t* rp, int32_t* row_of) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) row_of[e] = static_cast<int32_t>(r);
}

__global__ void iota_kernel(int64_t n, int32_t* v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = static_cast<int32_t>(i);
}

// t_rp[c] = first position k with sorted_key[k] >= c.
__global__ void offsets_from_sorted_kernel(int64_t nnz, int64_t cols, const int32_t* keys,
                                           int64_t* t_rp) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int64_t prev = (k == 0) ? -1 : static_cast<int64_t>(keys[k - 1]);
    const int64_t cur = (k == nnz) ? cols : static_cast<int64_t>(keys[k]);
    for (int64_t c = prev + 1; c <= cur; ++c) t_rp[c] = k;
}

template <typename V>
__global__ void transpose_gather_kernel(int64_t nnz, const int32_t* order, const int32_t* row_of,
                                        const V* val, int32_t* t_ci, V* t_val) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t e = order[k];
    t_ci[k] = row_of[e];
    t_val[k] = val[e];
}

// ------------------------------------------------------------ slab copy
// Sort key (panel * slabs + slab) and packed index of every entry; a thread
// per row finds its panel by bisection of the panel boundaries.
__global__ void slab_key_kernel(int64_t rows, const int64_t* rp, const int32_t* ci, const int64_t* panels,
                                int64_t npanels, int nslab, int32_t* key, uint32_t* packed) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    int64_t lo = 0, hi = npanels;  // panels[lo] <= r < panels[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (panels[mid] <= r)
            lo = mid;
        else
            hi = mid;
    }
    const int32_t kb = static_cast<int32_t>(lo * nslab);
    const uint32_t rl = static_cast<uint32_t>(r - panels[lo]) << 16;
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
        const int32_t c = ci[e];
        key[e] = kb + c / kSlabCols;
        packed[e] = rl | static_cast<uint32_t>(c % kSlabCols);
    }
}

template <typename V>
__global__ void slab_gather_kernel(int64_t nnz, const int32_t* order, const uint32_t* packed, const V* val,
                                   uint32_t* out_idx, V* out_val) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t e = order[k];
    out_idx[k] = packed[e];
    out_val[k] = val[e];
}

// ------------------------------------------------------------ stats
template <typename V>
__global__ __launch_bounds__(kThreads) void csr_row_stats_kernel(int64_t rows, const int64_t* rp,
                                                                 const V* val, double* sum, double* sq,
                                                                 double* misc_partial) {
    constexpr int G = 8;
    const int lig = threadIdx.x % G;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * (kThreads / G) + threadIdx.x / G;
    if (row >= rows) return;
    double s = 0.0, q = 0.0, mx = 0.0, bad = 0.0;
    for (int64_t e = rp[row] + lig; e < rp[row + 1]; e += G) {
        const double v = fabs(to_double(val[e]));
        if (!isfinite(v)) { bad += 1.0; continue; }
        s += v;
        q = fma(v, v, q);
        mx = fmax(mx, v);
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        s += __shfl_xor(s, off);
        q += __shfl_xor(q, off);
        mx = fmax(mx, __shfl_xor(mx, off));
        bad += __shfl_xor(bad, off);
    }
    if (lig == 0) {
        sum[row] = s;
        sq[row] = q;
        misc_partial[2 * row] = mx;
        misc_partial[2 * row + 1] = bad;
    }
}

__global__ void misc_reduce_kernel(int64_t rows, const double* misc_partial, double* misc) {
    __shared__ double smax[kThreads];
    __shared__ double scnt[kThreads];
    double m = 0.0, bad = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += kThreads) {
        m = fmax(m, misc_partial[2 * r]);
        bad += misc_partial[2 * r + 1];
    }
    smax[threadIdx.x] = m;
    scnt[threadIdx.x] = bad;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
            scnt[threadIdx.x] += scnt[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        misc[0] = smax[0];
        misc[1] = scnt[0];
    }
}

__global__ void csr_validate_kernel(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rp,
                                    const int32_t* ci, int* flags) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r == 0 && (rp[0] != 0 || rp[rows] != nnz)) flags[0] = 0;
    if (r >= rows) return;
    const int64_t e0 = rp[r], e1 = rp[r + 1];
    if (e1 < e0 || e0 < 0 || e1 > nnz) {
        flags[0] = 0;
        return;
    }
    bool sorted = true;
    for (int64_t e = e0; e < e1; ++e) {
        const int32_t c = ci[e];
        if (c < 0 || c >= cols) flags[1] = 0;
        if (e + 1 < e1 && ci[e + 1] <= c) sorted = false;
    }
    if (!sorted) {
        for (int64_t e = e0; e < e1; ++e)
            for (int64_t f = e + 1; f < e1; ++f)
                if (ci[e] == ci[f]) flags[2] = 0;
    }
}

}  // namespace

// The stream and slab kernels follow the table's variants.
constexpr int kStreamVariant = kVariantCount;
constexpr int kStreamWideVariant = kVariantCount + 1;
constexpr int kSlabVariant = kVariantCount + 2;

int csr_spmm_variant_count() { return kVariantCount + 3; }
int csr_spmm_stream_variant() { return kStreamVariant; }
int csr_stream_budget() { return kStreamBudget; }
int csr_spmm_slab_variant() { return kSlabVariant; }
int csr_slab_cols() { return kSlabCols; }

int csr_slab_count(int64_t cols) {
    const int64_t n = ceil_div(cols, static_cast<int64_t>(kSlabCols));
    return n <= kSlabMaxCount ? static_cast<int>(n) : -1;
}

int64_t csr_slab_panel_entries(int64_t nnz) {
    // about kSlabPanels panels, each staging every slab of x it touches
    return std::max<int64_t>(4096, ceil_div(nnz, static_cast<int64_t>(kSlabPanels)));
}

std::vector<int64_t> csr_slab_partition(const std::vector<int64_t>& rp, int64_t panel_entries) {
    const int64_t rows = static_cast<int64_t>(rp.size()) - 1;
    std::vector<int64_t> br{0};
    int64_t r = 0;
    while (r < rows) {
        const int64_t e0 = rp[r];
        int64_t end = r + 1;
        while (end < rows && rp[end + 1] - e0 <= panel_entries && end - r < kSlabRows) ++end;
        br.push_back(end);
        r = end;
    }
    return br;
}

void csr_slab_build(hipStream_t s, const CsrView& A, const int64_t* panels, int64_t npanels, int slab_count,
                    int64_t* tiles, uint32_t* idx, void* val, Scratch& scratch) {
    const int64_t nnz = A.nnz;
    const int64_t ntiles = npanels * slab_count;
    if (slab_count < 1 || npanels < 1 || nnz >= (int64_t(1) << 31) || ntiles >= (int64_t(1) << 31) - 1)
        throw std::invalid_argument("CSR slab copy supports nnz and tiles below 2^31");
    if (nnz == 0) {
        AMFLS_HIP_CHECK(hipMemsetAsync(tiles, 0, static_cast<size_t>(ntiles + 1) * sizeof(int64_t), s));
        return;
    }
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < ntiles) ++end_bit;
    size_t temp_bytes = 0;
    AMFLS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, static_cast<const int32_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr),
                                                       static_cast<const int32_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr), static_cast<int>(nnz), 0,
                                                       end_bit, s));
    const size_t arr = static_cast<size_t>(nnz) * sizeof(int32_t);
    const size_t align = 256;
    const size_t arr_al = (arr + align - 1) / align * align;
    scratch.ensure(5 * arr_al + temp_bytes + align);
    char* base = static_cast<char*>(scratch.ptr);
    int32_t* keys_in = reinterpret_cast<int32_t*>(base);
    int32_t* keys_out = reinterpret_cast<int32_t*>(base + arr_al);
    int32_t* vals_in = reinterpret_cast<int32_t*>(base + 2 * arr_al);
    int32_t* vals_out = reinterpret_cast<int32_t*>(base + 3 * arr_al);
    uint32_t* packed = reinterpret_cast<uint32_t*>(base + 4 * arr_al);
    void* temp = base + 5 * arr_al;
    hipLaunchKernelGGL(slab_key_kernel, dim3(static_cast<unsigned>(ceil_div(A.rows, kThreads))), dim3(kThreads), 0,
                       s, A.rows, A.row_ptr, A.col_idx, panels, npanels, slab_count, keys_in, packed);
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(iota_kernel, dim3(static_cast<unsigned>(ceil_div(nnz, kThreads))), dim3(kThreads), 0, s,
                       nnz, vals_in);
    AMFLS_LAUNCH_CHECK();
    AMFLS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                       static_cast<int>(nnz), 0, end_bit, s));
    hipLaunchKernelGGL(offsets_from_sorted_kernel, dim3(static_cast<unsigned>(ceil_div(nnz + 1, kThreads))),
                       dim3(kThreads), 0, s, nnz, ntiles, keys_out, tiles);
    AMFLS_LAUNCH_CHECK();
    const dim3 g(static_cast<unsigned>(ceil_div(nnz, kThreads)));
    if (A.type == ValueType::f32)
        hipLaunchKernelGGL(slab_gather_kernel<float>, g, dim3(kThreads), 0, s, nnz, vals_out, packed,
                           static_cast<const float*>(A.values), idx, static_cast<float*>(val));
    else
        hipLaunchKernelGGL(slab_gather_kernel<double>, g, dim3(kThreads), 0, s, nnz, vals_out, packed,
                           static_cast<const double*>(A.values), idx, static_cast<double*>(val));
    AMFLS_LAUNCH_CHECK();
}

std::vector<int64_t> csr_stream_partition(const std::vector<int64_t>& rp, bool wide) {
    const int64_t rows = static_cast<int64_t>(rp.size()) - 1;
    const int64_t budget = wide ? kStreamWideBudget : kStreamBudget;
    const int64_t max_rows = wide ? kStreamWideMaxRows : kStreamMaxRows;
    std::vector<int64_t> br{0};
    int64_t r = 0;
    while (r < rows) {
        const int64_t e0 = rp[r];
        int64_t end = r + 1;
        while (end < rows && rp[end + 1] - e0 <= budget && end - r < max_rows) ++end;
        br.push_back(end);
        r = end;
    }
    return br;
}

bool csr_spmm_variant_supports(int variant, int ncols) {
    if (variant == kStreamVariant || variant == kStreamWideVariant || variant == kSlabVariant) return ncols == 1;
    if (variant < 0 || variant >= kVariantCount) return false;
    return ncols <= kVariants[variant].lc * kVariants[variant].cpl;
}

const char* csr_spmm_variant_name(int variant) {
    if (variant == kStreamVariant) return "stream";
    if (variant == kStreamWideVariant) return "stream8k";
    if (variant == kSlabVariant) return "slab";
    return (variant >= 0 && variant < kVariantCount) ? kVariants[variant].name : "invalid";
}

bool csr_spmm_variant_shape(int variant, int* p, int* lc, int* cpl) {
    if (variant == kStreamVariant || variant == kStreamWideVariant || variant == kSlabVariant) {
        *p = 0;
        *lc = 1;
        *cpl = 1;
        return true;
    }
    if (variant < 0 || variant >= kVariantCount) return false;
    *p = kVariants[variant].p;
    *lc = kVariants[variant].lc;
    *cpl = kVariants[variant].cpl;
    return true;
}

int csr_spmm_variant_index(int p, int lc, int cpl) {
    for (int v = 0; v < kVariantCount; ++v)
        if (kVariants[v].p == p && kVariants[v].lc == lc && kVariants[v].cpl == cpl) return v;
    return -1;
}

int csr_spmm_default_variant(int ncols, double mean_row_nnz) {
    // Column lanes cover the block width (up to one wavefront); entry lanes
    // split a row so that each keeps about three or more entries per column
    // lane (measured on gfx950 for rows of 12 and 120 entries, widths 1-64).
    if (ncols > 128) return csr_spmm_variant_index(1, 64, 4);
    if (ncols > 64) return csr_spmm_variant_index(1, 64, 2);
    int lc = 1;
    while (lc < ncols) lc <<= 1;
    int p = 1;
    while (p * lc < 64 && 3.0 * lc * (2 * p) <= mean_row_nnz + 1e-9) p <<= 1;
    return csr_spmm_variant_index(p, lc, 1);
}

template <typename T>
void csr_spmm(hipStream_t s, const CsrView& A, int ncols, const T* X, int64_t ldx, T* Y, int64_t ldy,
              const T* Add, int64_t ld_add, double add_scale, int variant, hipEvent_t ev_start,
              hipEvent_t ev_stop, unsigned long long* stamp) {
    if (A.rows <= 0 || ncols <= 0) {
        // keep a caller's event pair well defined for an empty product
        if (ev_start) AMFLS_HIP_CHECK(hipEventRecord(ev_start, s));
        if (ev_stop) AMFLS_HIP_CHECK(hipEventRecord(ev_stop, s));
        // and a zero start slot, so the timer charges an empty product nothing
        // (the slot may still hold an earlier product's stamps)
        if (stamp) AMFLS_HIP_CHECK(hipMemsetAsync(stamp, 0, sizeof(unsigned long long), s));
        return;
    }
    // Blocks wider than 256 columns are processed in 256-column slabs.
    for (int c0 = 0; c0 < ncols; c0 += 256) {
        const int nc = std::min(256, ncols - c0);
        int v = variant;
        if ((v == kStreamVariant && A.stream_rows == nullptr) ||
            (v == kStreamWideVariant && A.stream_rows_wide == nullptr) ||
            (v == kSlabVariant && A.slab_panels == nullptr))
            v = -1;
        if (v < 0 || !csr_spmm_variant_supports(v, nc))
            v = csr_spmm_default_variant(nc, A.rows > 0 ? static_cast<double>(A.nnz) / A.rows : 0.0);
        const T* add = Add ? Add + c0 : nullptr;
        hipEvent_t e0 = c0 == 0 ? ev_start : nullptr;
        hipEvent_t e1 = c0 + nc == ncols ? ev_stop : nullptr;
        const ProductStamp st{stamp, c0 == 0, c0 + nc == ncols};
        if (v == kStreamVariant) {
            launch_stream<T, kStreamBudget>(s, A, X, ldx, Y, ldy, add, ld_add, add_scale, e0, e1, st);
            continue;
        }
        if (v == kStreamWideVariant) {
            launch_stream<T, kStreamWideBudget>(s, A, X, ldx, Y, ldy, add, ld_add, add_scale, e0, e1, st);
            continue;
        }
        if (v == kSlabVariant) {
            launch_slab<T>(s, A, X, ldx, Y, ldy, add, ld_add, add_scale, e0, e1, st);
            continue;
        }
        switch (A.type) {
            case ValueType::f32:
                launch_by_index<float, T>(v, s, A, nc, X + c0, ldx, Y + c0, ldy, add, ld_add, add_scale, e0, e1,
                                          st);
                break;
            case ValueType::f64:
                launch_by_index<double, T>(v, s, A, nc, X + c0, ldx, Y + c0, ldy, add, ld_add, add_scale, e0, e1,
                                           st);
                break;
            default:
                throw std::invalid_argument("CSR values must be float32 or float64");
        }
    }
}

void csr_transpose(hipStream_t s, const CsrView& A, int64_t* t_rp, int32_t* t_ci, void* t_val,
                   Scratch& scratch) {
    const int64_t nnz = A.nnz;
    if (nnz >= (int64_t(1) << 31)) throw std::invalid_argument("CSR transpose supports nnz < 2^31");
    if (nnz == 0) {
        AMFLS_HIP_CHECK(hipMemsetAsync(t_rp, 0, static_cast<size_t>(A.cols + 1) * sizeof(int64_t), s));
        return;
    }
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < A.cols) ++end_bit;
    size_t temp_bytes = 0;
    AMFLS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, static_cast<const int32_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr),
                                                       static_cast<const int32_t*>(nullptr),
                                                       static_cast<int32_t*>(nullptr), static_cast<int>(nnz), 0,
                                                       end_bit, s));
    const size_t arr = static_cast<size_t>(nnz) * sizeof(int32_t);
    const size_t align = 256;
    const size_t arr_al = (arr + align - 1) / align * align;
    scratch.ensure(4 * arr_al + temp_bytes + align);
    char* base = static_cast<char*>(scratch.ptr);
    int32_t* keys_out = reinterpret_cast<int32_t*>(base);
    int32_t* vals_in = reinterpret_cast<int32_t*>(base + arr_al);
    int32_t* vals_out = reinterpret_cast<int32_t*>(base + 2 * arr_al);
    int32_t* row_of = reinterpret_cast<int32_t*>(base + 3 * arr_al);
    void* temp = base + 4 * arr_al;
    hipLaunchKernelGGL(row_of_kernel, dim3(static_cast<unsigned>(ceil_div(A.rows, kThreads))), dim3(kThreads), 0,
                       s, A.rows, A.row_ptr, row_of);
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(iota_kernel, dim3(static_cast<unsigned>(ceil_div(nnz, kThreads))), dim3(kThreads), 0, s,
                       nnz, vals_in);
    AMFLS_LAUNCH_CHECK();
    AMFLS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, A.col_idx, keys_out, vals_in,
                                                       vals_out, static_cast<int>(nnz), 0, end_bit, s));
    hipLaunchKernelGGL(offsets_from_sorted_kernel, dim3(static_cast<unsigned>(ceil_div(nnz + 1, kThreads))),
                       dim3(kThreads), 0, s, nnz, A.cols, keys_out, t_rp);
    AMFLS_LAUNCH_CHECK();
    const dim3 g(static_cast<unsigned>(ceil_div(nnz, kThreads)));
    if (A.type == ValueType::f32)
        hipLaunchKernelGGL(transpose_gather_kernel<float>, g, dim3(kThreads), 0, s, nnz, vals_out, row_of,
                           static_cast<const float*>(A.values), t_ci, static_cast<float*>(t_val));
    else
        hipLaunchKernelGGL(transpose_gather_kernel<double>, g, dim3(kThreads), 0, s, nnz, vals_out, row_of,
                           static_cast<const double*>(A.values), t_ci, static_cast<double*>(t_val));
    AMFLS_LAUNCH_CHECK();
}

void csr_row_stats(hipStream_t s, const CsrView& A, double* out_sum, double* out_sq, double* misc,
                   Scratch& scratch) {
    if (A.rows <= 0) return;
    scratch.ensure(static_cast<size_t>(2 * A.rows) * sizeof(double));
    double* partial = static_cast<double*>(scratch.ptr);
    const dim3 g(static_cast<unsigned>(ceil_div(A.rows, kThreads / 8)));
    if (A.type == ValueType::f32)
        hipLaunchKernelGGL(csr_row_stats_kernel<float>, g, dim3(kThreads), 0, s, A.rows, A.row_ptr,
                           static_cast<const float*>(A.values), out_sum, out_sq, partial);
    else
        hipLaunchKernelGGL(csr_row_stats_kernel<double>, g, dim3(kThreads), 0, s, A.rows, A.row_ptr,
                           static_cast<const double*>(A.values), out_sum, out_sq, partial);
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(misc_reduce_kernel, dim3(1), dim3(kThreads), 0, s, A.rows, partial, misc);
    AMFLS_LAUNCH_CHECK();
}

void csr_validate(hipStream_t s, const CsrView& A, int* out_dev) {
    fill<int>(s, out_dev, 3, 1);
    hipLaunchKernelGGL(csr_validate_kernel, dim3(static_cast<unsigned>(ceil_div(A.rows, kThreads))),
                       dim3(kThreads), 0, s, A.rows, A.cols, A.nnz, A.row_ptr, A.col_idx, out_dev);
    AMFLS_LAUNCH_CHECK();
}

template void csr_spmm<float>(hipStream_t, const CsrView&, int, const float*, int64_t, float*, int64_t,
                              const float*, int64_t, double, int, hipEvent_t, hipEvent_t,
                              unsigned long long*);
template void csr_spmm<double>(hipStream_t, const CsrView&, int, const double*, int64_t, double*,
                               int64_t, const double*, int64_t, double, int, hipEvent_t, hipEvent_t,
                               unsigned long long*);

}  // namespace amflsx
