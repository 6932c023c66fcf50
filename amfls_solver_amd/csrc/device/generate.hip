// Counter-addressed synthetic CSR matrices on the device.
//
// Every quantity is a pure function of (seed, stream, counter), built from
// the splitmix64 / open-uniform / Box-Muller chain of the reference's
// gaussian_at (src/algorithms/random/gaussian.cpp:13-35):
//   stream 1, counter r            row length  lo + floor(u * (hi - lo + 1))
//   stream 2, counter r * 2^16 + a column draw a of row r: floor(u * cols);
//                                  rejection of repeats, at most 2^16 draws,
//                                  then the next unused column upward
//   stream 3, counter e            value of entry e (row-sorted order): N(0,1)
//   stream 4, counter r            row scale 10^(l0 + u * (l1 - l0))
// Columns of a row are sorted ascending.  The numpy restatement in the
// oracle reproduces the same integers and, up to libm rounding, the values.
#include "kernels.hpp"

#include <hipcub/hipcub.hpp>

namespace amflsx {
namespace {

constexpr int kGenThreads = 256;
constexpr int kMaxRow = 64;

__device__ __forceinline__ uint64_t mix(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

__device__ __forceinline__ double unit(uint64_t bits) {
    return static_cast<double>((bits >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ int64_t draw_below(uint64_t key, uint64_t counter, int64_t bound) {
    const int64_t v = static_cast<int64_t>(unit(mix(key ^ counter)) * static_cast<double>(bound));
    return v < bound ? v : bound - 1;  // u = 1 maps to the last value
}

__global__ void row_length_kernel(int64_t rows, int64_t cols, int lo, int hi, uint64_t key, int64_t* len) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r > rows) return;
    if (r == rows) {
        len[r] = 0;
        return;
    }
    const int64_t k = lo + draw_below(key, static_cast<uint64_t>(r), hi - lo + 1);
    len[r] = k < cols ? k : cols;
}

template <typename V>
__global__ __launch_bounds__(kGenThreads) void entries_kernel(int64_t rows, int64_t cols, const int64_t* row_ptr,
                                                              uint64_t key_col, uint64_t key_val, uint64_t key_scale,
                                                              double log_lo, double log_hi, int32_t* col_idx,
                                                              V* values) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t e0 = row_ptr[r];
    const int k = static_cast<int>(row_ptr[r + 1] - e0);
    int32_t c[kMaxRow];
    int have = 0;
    uint32_t attempt = 0;
    while (have < k) {
        int64_t v;
        bool fresh;
        if (attempt < 65536u) {
            v = draw_below(key_col, (static_cast<uint64_t>(r) << 16) + attempt, cols);
            ++attempt;
            fresh = true;
            for (int i = 0; i < have; ++i) fresh = fresh && c[i] != v;
        } else {
            // after 2^16 draws: the smallest unused column (k <= cols, so one exists)
            v = 0;
            for (bool hit = true; hit;) {
                hit = false;
                for (int i = 0; i < have; ++i) hit = hit || c[i] == v;
                if (hit) ++v;
            }
            fresh = true;
        }
        if (!fresh) continue;
        int pos = have++;  // insertion keeps c sorted
        while (pos > 0 && c[pos - 1] > v) {
            c[pos] = c[pos - 1];
            --pos;
        }
        c[pos] = static_cast<int32_t>(v);
    }
    const double scale = (log_lo == 0.0 && log_hi == 0.0)
                             ? 1.0
                             : exp10(log_lo + unit(mix(key_scale ^ static_cast<uint64_t>(r))) * (log_hi - log_lo));
    for (int i = 0; i < k; ++i) {
        const uint64_t e = static_cast<uint64_t>(e0 + i);
        const double u1 = unit(mix(key_val ^ (2ULL * e)));
        const double u2 = unit(mix(key_val ^ (2ULL * e + 1ULL)));
        const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
        col_idx[e0 + i] = c[i];
        values[e0 + i] = static_cast<V>(z * scale);
    }
}

uint64_t host_mix(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

uint64_t key_of(uint64_t seed, uint64_t stream) { return host_mix(seed) ^ host_mix(stream + 0xd2b74407b1ce6e93ULL); }

}  // namespace

int64_t csr_generate_rows(hipStream_t s, int64_t rows, int64_t cols, int lo, int hi, uint64_t seed,
                          int64_t* row_ptr, Scratch& scratch) {
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("generated CSR needs positive dimensions");
    if (lo < 0 || hi < lo || hi > kMaxRow) throw std::invalid_argument("row lengths must satisfy 0 <= lo <= hi <= 64");
    const size_t len_bytes = static_cast<size_t>(rows + 1) * sizeof(int64_t);
    size_t temp = 0;
    AMFLS_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, static_cast<const int64_t*>(nullptr),
                                                     static_cast<int64_t*>(nullptr), static_cast<int>(rows + 1), s));
    const size_t aligned = (len_bytes + 255) / 256 * 256;
    scratch.ensure(aligned + temp + 256);
    int64_t* len = static_cast<int64_t*>(scratch.ptr);
    void* tmp = static_cast<char*>(scratch.ptr) + aligned;
    if (rows + 1 > (int64_t(1) << 31) - 1) throw std::invalid_argument("generated CSR supports rows < 2^31 - 1");
    hipLaunchKernelGGL(row_length_kernel, dim3(static_cast<unsigned>(ceil_div(rows + 1, kGenThreads))),
                       dim3(kGenThreads), 0, s, rows, cols, lo, hi, key_of(seed, 1), len);
    AMFLS_LAUNCH_CHECK();
    AMFLS_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, temp, len, row_ptr, static_cast<int>(rows + 1), s));
    int64_t nnz = 0;
    AMFLS_HIP_CHECK(hipMemcpyAsync(&nnz, row_ptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AMFLS_HIP_CHECK(hipStreamSynchronize(s));
    return nnz;
}

void csr_generate_entries(hipStream_t s, int64_t rows, int64_t cols, uint64_t seed, double scale_lo,
                          double scale_hi, const int64_t* row_ptr, int32_t* col_idx, void* values, ValueType type) {
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("generated CSR needs positive dimensions");
    if (cols > (int64_t(1) << 31) - 1) throw std::invalid_argument("generated CSR supports cols < 2^31");
    if (!(scale_lo > 0.0) || !(scale_hi >= scale_lo) || !std::isfinite(scale_hi))
        throw std::invalid_argument("row scales must satisfy 0 < lo <= hi < inf");
    const double l0 = std::log10(scale_lo), l1 = std::log10(scale_hi);
    const dim3 grid(static_cast<unsigned>(ceil_div(rows, kGenThreads)));
    const uint64_t kc = key_of(seed, 2), kv = key_of(seed, 3), ks = key_of(seed, 4);
    switch (type) {
        case ValueType::f32:
            hipLaunchKernelGGL(entries_kernel<float>, grid, dim3(kGenThreads), 0, s, rows, cols, row_ptr, kc, kv, ks,
                               l0, l1, col_idx, static_cast<float*>(values));
            break;
        case ValueType::f64:
            hipLaunchKernelGGL(entries_kernel<double>, grid, dim3(kGenThreads), 0, s, rows, cols, row_ptr, kc, kv,
                               ks, l0, l1, col_idx, static_cast<double*>(values));
            break;
        default:
            throw std::invalid_argument("generated CSR values are float32 or float64");
    }
    AMFLS_LAUNCH_CHECK();
}

}  // namespace amflsx
