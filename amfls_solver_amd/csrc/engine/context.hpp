// Per-solve device context: one HIP stream, scratch arenas, small host<->device
// transfers and the (wall or synthetic) clock used by the width controller.
#pragma once

#include <chrono>
#include <cstring>
#include <utility>
#include <vector>

#include "../device/kernels.hpp"
#include "types.hpp"

namespace amflsx {

template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { reset(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void reset(size_t n) {
        release();
        if (n > 0) AMFLS_HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
        n_ = n;
    }
    void ensure(size_t n) {
        if (n > n_) reset(n);
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

class Clock {
public:
    void configure(ClockMode mode, double op_unit, double local_unit) {
        synthetic_ = mode == ClockMode::synthetic;
        op_unit_ = op_unit;
        local_unit_ = local_unit;
        t_ = 0.0;
    }
    bool synthetic() const { return synthetic_; }
    double now() const {
        if (synthetic_) return t_;
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    void charge_operator(double cost) {
        if (synthetic_) t_ += op_unit_ * cost;
    }
    void charge_local(int block_cols, int basis_cols) {
        if (synthetic_) t_ += local_unit_ * static_cast<double>(block_cols) *
                              static_cast<double>(basis_cols + block_cols);
    }

private:
    bool synthetic_ = false;
    double op_unit_ = 1e-3, local_unit_ = 1e-6, t_ = 0.0;
};

struct Context {
    hipStream_t stream = nullptr;
    Scratch scratch;   // kernel-internal partials
    Scratch scratch2;  // a second arena for nested needs (transpose build)
    DevBuf<double> small;  // device staging for small matrices (coefficients, R^-1)
    DevBuf<double> stats;  // device results of reductions
    Clock clock;

    Context() {
        AMFLS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        stats.reset(64);
    }
    ~Context() {
        scratch.release();
        scratch2.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    void sync() { AMFLS_HIP_CHECK(hipStreamSynchronize(stream)); }

    // Upload a small double matrix into the staging buffer at `offset`
    // doubles; returns the device pointer.  Synchronous with respect to the
    // host buffer (the copy is complete when this returns).
    double* upload(const double* host, size_t count, size_t offset = 0) {
        small.ensure(offset + count + 1);
        double* dst = small.get() + offset;
        if (count > 0) {
            AMFLS_HIP_CHECK(hipMemcpyAsync(dst, host, count * sizeof(double), hipMemcpyHostToDevice, stream));
            sync();
        }
        return dst;
    }
    template <typename T>
    void download(T* host, const T* dev, size_t count) {
        if (count == 0) return;
        AMFLS_HIP_CHECK(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, stream));
        sync();
    }

    // Statistics (max |v|, scaled square sum, nonfinite count) of a block.
    template <typename T>
    void block_stats(int64_t rows, int cols, const T* X, int64_t ldx, double out[3]) {
        vector_stats<T>(stream, rows, cols, X, ldx, stats.get(), scratch);
        download(out, stats.get(), 3);
    }
};

}  // namespace amflsx
