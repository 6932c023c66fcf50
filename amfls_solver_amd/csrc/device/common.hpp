// Shared device/host helpers for the AMFLS HIP engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace amflsx {

class HipError : public std::runtime_error {
public:
    explicit HipError(const std::string& what) : std::runtime_error(what) {}
};

#define AMFLS_HIP_CHECK(expr)                                                   \
    do {                                                                        \
        hipError_t amfls_status_ = (expr);                                      \
        if (amfls_status_ != hipSuccess) {                                      \
            throw ::amflsx::HipError(std::string(#expr) + " failed: " +       \
                                     hipGetErrorString(amfls_status_));         \
        }                                                                       \
    } while (0)

#define AMFLS_LAUNCH_CHECK() AMFLS_HIP_CHECK(hipGetLastError())

constexpr int kWave = 64;

// Row-major strided block: element (i, j) lives at ptr[i * ld + j].
template <typename T>
struct Block {
    T* ptr = nullptr;
    int64_t rows = 0;
    int cols = 0;
    int64_t ld = 0;
    __host__ __device__ T* row(int64_t i) const { return ptr + i * ld; }
};

// Value-type conversion for stored matrix entries (bf16 / fp32 / fp64).
__device__ __forceinline__ float to_float(float v) { return v; }
__device__ __forceinline__ float to_float(double v) { return static_cast<float>(v); }
__device__ __forceinline__ float to_float(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ double to_double(float v) { return static_cast<double>(v); }
__device__ __forceinline__ double to_double(double v) { return v; }
__device__ __forceinline__ double to_double(__hip_bfloat16 v) {
    return static_cast<double>(__bfloat162float(v));
}

template <typename T> __device__ __forceinline__ T from_double(double v);
template <> __device__ __forceinline__ double from_double<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_double<float>(double v) {
    return static_cast<float>(v);
}

template <typename Acc, typename V> __device__ __forceinline__ Acc load_as(V v);
template <> __device__ __forceinline__ double load_as<double, double>(double v) { return v; }
template <> __device__ __forceinline__ double load_as<double, float>(float v) { return v; }
template <> __device__ __forceinline__ double load_as<double, __hip_bfloat16>(__hip_bfloat16 v) {
    return to_double(v);
}
template <> __device__ __forceinline__ float load_as<float, float>(float v) { return v; }
template <> __device__ __forceinline__ float load_as<float, double>(double v) {
    return static_cast<float>(v);
}
template <> __device__ __forceinline__ float load_as<float, __hip_bfloat16>(__hip_bfloat16 v) {
    return __bfloat162float(v);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace amflsx
