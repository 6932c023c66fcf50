// Per-solve device context: one HIP stream, scratch arenas, small host<->device
// transfers and the (wall or synthetic) clock used by the width controller.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstring>
#include <stdexcept>
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
    // Grows geometrically; contents are not preserved.
    void ensure(size_t n) {
        if (n > n_) reset(std::max(n, 2 * n_));
    }
    // hipFree waits for the device, so a buffer still read by queued
    // kernels is never released early.
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

// Page-locked host memory (hipHostMalloc) for asynchronous small transfers.
template <typename T>
class PinnedBuf {
public:
    PinnedBuf() = default;
    ~PinnedBuf() { release(); }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    void ensure(size_t n) {
        if (n <= n_) return;
        release();
        AMFLS_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
        n_ = n;
    }
    void release() {
        if (p_) (void)hipHostFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// GPU-side timing of operator products: an event pair around each product,
// resolved (and added to its target counter) on flush().  Nothing waits on
// the device between products; flush() waits for the last recorded event.
//
// Operators whose product kernels time themselves (ProductStamp) skip the
// events: kernel_stamp() hands out one product's stamp values in device
// memory and flush() reads them back.  Event records on the stream put a gap
// of several microseconds before and after each product; in-kernel stamps
// add none.
class ProductTimer {
public:
    ~ProductTimer() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        if (dstamp_) (void)hipFree(dstamp_);
        if (hstamp_) (void)hipHostFree(hstamp_);
    }
    // The kStampSlots device values that the next product's kernels stamp;
    // the product's time is added to *target (and *target2, if given) on
    // flush().
    unsigned long long* kernel_stamp(hipStream_t s, double* target, double* target2 = nullptr) {
        if (dstamp_ == nullptr) {
            tick_seconds_ = wall_tick_seconds();
            allocate(s);
        }
        if (!stamp_targets_.empty() && s != stamp_stream_) flush();
        if (stamp_targets_.size() == static_cast<size_t>(kStampStride)) flush();  // reuse from product 0
        stamp_stream_ = s;
        unsigned long long* slots = dstamp_ + stamp_targets_.size();
        stamp_targets_.push_back({target, target2});
        return slots;
    }
    // Make room for the stamps of `products` further products on `s` now
    // (allocating or flushing here), so that handing them out never
    // synchronizes a stream: for products enqueued under stream capture.
    void reserve(hipStream_t s, size_t products) {
        if (products > static_cast<size_t>(kStampStride))
            throw std::length_error("more stamped products than the stamp buffer holds");
        if (dstamp_ == nullptr) {
            tick_seconds_ = wall_tick_seconds();
            allocate(s);
        }
        if (!stamp_targets_.empty() && (s != stamp_stream_ || stamp_targets_.size() + products >
                                                                  static_cast<size_t>(kStampStride)))
            flush();
    }
    void begin(hipStream_t s) {
        open_ = take();
        AMFLS_HIP_CHECK(hipEventRecord(open_, s));
    }
    void end(hipStream_t s, double* target) {
        hipEvent_t stop = take();
        AMFLS_HIP_CHECK(hipEventRecord(stop, s));
        pending_.push_back({open_, stop, target});
        open_ = nullptr;
    }
    // Drop unresolved pairs (their targets may no longer exist).
    void discard() {
        for (const auto& p : pending_) {
            pool_.push_back(p.start);
            pool_.push_back(p.stop);
        }
        pending_.clear();
        if (open_) pool_.push_back(open_);
        open_ = nullptr;
        stamp_targets_.clear();
    }
    void flush() {
        if (!stamp_targets_.empty()) {
            const size_t n = stamp_targets_.size();
            // No clearing between uses: every launched product rewrites slot 0,
            // and end slots left from an earlier product hold earlier clock
            // values, below this product's own ends, so the maximum is unchanged.
            // Slot j of the n pending products: n contiguous values per row.
            AMFLS_HIP_CHECK(hipMemcpy2DAsync(hstamp_, n * sizeof(unsigned long long), dstamp_,
                                             kStampStride * sizeof(unsigned long long),
                                             n * sizeof(unsigned long long), kStampSlots, hipMemcpyDeviceToHost,
                                             stamp_stream_));
            AMFLS_HIP_CHECK(hipStreamSynchronize(stamp_stream_));
            for (size_t i = 0; i < n; ++i) {
                unsigned long long v[kStampSlots];
                for (int j = 0; j < kStampSlots; ++j) v[j] = hstamp_[j * n + i];
                unsigned long long end = 0;
                for (int j = 1; j < kStampSlots; ++j) end = std::max(end, v[j]);
                if (v[0] == 0 || end <= v[0]) continue;
                const double sec = tick_seconds_ * static_cast<double>(end - v[0]);
                *stamp_targets_[i].first += sec;
                if (stamp_targets_[i].second) *stamp_targets_[i].second += sec;
            }
            stamp_targets_.clear();
        }
        if (pending_.empty()) return;
        AMFLS_HIP_CHECK(hipEventSynchronize(pending_.back().stop));
        for (const auto& p : pending_) {
            float ms = 0.0f;
            AMFLS_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
            *p.target += 1e-3 * static_cast<double>(ms);
            pool_.push_back(p.start);
            pool_.push_back(p.stop);
        }
        pending_.clear();
    }

private:
    struct Pending { hipEvent_t start, stop; double* target; };
    hipEvent_t take() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        AMFLS_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    std::vector<hipEvent_t> pool_;
    std::vector<Pending> pending_;
    hipEvent_t open_ = nullptr;
    static double wall_tick_seconds() {
        int dev = 0, khz = 0;
        AMFLS_HIP_CHECK(hipGetDevice(&dev));
        AMFLS_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) throw std::runtime_error("device reports no wall clock rate");
        return 1.0 / (1e3 * static_cast<double>(khz));
    }
    // Stamps of kStampStride products, slot-major (2.1 MB on the device).
    void allocate(hipStream_t s) {
        const size_t bytes = static_cast<size_t>(kStampStride) * kStampSlots * sizeof(unsigned long long);
        AMFLS_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dstamp_), bytes));
        AMFLS_HIP_CHECK(hipMemsetAsync(dstamp_, 0, bytes, s));
        AMFLS_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hstamp_), bytes, hipHostMallocDefault));
    }
    unsigned long long* dstamp_ = nullptr;
    unsigned long long* hstamp_ = nullptr;
    hipStream_t stamp_stream_ = nullptr;
    std::vector<std::pair<double*, double*>> stamp_targets_;
    double tick_seconds_ = 0.0;
};

struct Context {
    hipStream_t stream = nullptr;
    Scratch scratch;   // kernel-internal partials
    Scratch scratch2;  // a second arena for nested needs (transpose build)
    DevBuf<double> small;  // device staging for small matrices (coefficients, R^-1)
    DevBuf<double> stats;  // device results of reductions
    DevBuf<unsigned> counter;  // ticket word of one-launch reductions (0 between launches)
    PinnedBuf<double> stage;   // host side of asynchronous uploads
    size_t stage_pos = 0;
    PinnedBuf<char> landing;   // host side of small downloads
    ProductTimer timer;
    Clock clock;
    // Width-one level batches are captured into a HIP graph and launched as
    // one graph instead of being enqueued kernel by kernel.
    bool level_graphs = false;

    static constexpr size_t kStageDoubles = 1 << 20;
    static constexpr size_t kLandingBytes = 1 << 20;

    Context() {
        AMFLS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        stats.reset(64);
        counter.reset(1);
        AMFLS_HIP_CHECK(hipMemset(counter.get(), 0, sizeof(unsigned)));
        stage.ensure(kStageDoubles);
        landing.ensure(kLandingBytes);
    }
    ~Context() {
        if (stream) (void)hipStreamSynchronize(stream);
        scratch.release();
        scratch2.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    void sync() {
        AMFLS_HIP_CHECK(hipStreamSynchronize(stream));
        stage_pos = 0;
    }

    // Upload a small double matrix into the device staging buffer at
    // `offset` doubles and return the device pointer.  The copy is enqueued
    // on the stream from page-locked memory and the host buffer may be
    // reused at once; kernels enqueued afterwards see the data.
    double* upload(const double* host, size_t count, size_t offset = 0) {
        small.ensure(offset + count + 1);
        double* dst = small.get() + offset;
        if (count == 0) return dst;
        if (count > kStageDoubles) {
            AMFLS_HIP_CHECK(hipMemcpyAsync(dst, host, count * sizeof(double), hipMemcpyHostToDevice, stream));
            sync();
            return dst;
        }
        if (stage_pos + count > kStageDoubles) sync();
        double* h = stage.get() + stage_pos;
        std::memcpy(h, host, count * sizeof(double));
        AMFLS_HIP_CHECK(hipMemcpyAsync(dst, h, count * sizeof(double), hipMemcpyHostToDevice, stream));
        stage_pos += count;
        return dst;
    }
    // Copy device data to the host and wait for it.
    template <typename T>
    void download(T* host, const T* dev, size_t count) {
        if (count == 0) return;
        const size_t bytes = count * sizeof(T);
        if (bytes <= kLandingBytes) {
            AMFLS_HIP_CHECK(hipMemcpyAsync(landing.get(), dev, bytes, hipMemcpyDeviceToHost, stream));
            sync();
            std::memcpy(host, landing.get(), bytes);
            return;
        }
        AMFLS_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream));
        sync();
    }

    // Statistics (max |v|, scaled square sum, nonfinite count) of a block,
    // left on the device at `out_dev` (no wait).
    template <typename T>
    void block_stats_async(int64_t rows, int cols, const T* X, int64_t ldx, double* out_dev,
                           const double* gate = nullptr) {
        vector_stats<T>(stream, rows, cols, X, ldx, out_dev, scratch, counter.get(), gate);
    }
    // Statistics copied to the host (waits).
    template <typename T>
    void block_stats(int64_t rows, int cols, const T* X, int64_t ldx, double out[3]) {
        block_stats_async<T>(rows, cols, X, ldx, stats.get());
        download(out, stats.get(), 3);
    }
};

}  // namespace amflsx
