// Allocation-free scratch: one file-scope device array per GPU, for the
// product entry point that callers who cannot allocate device memory use
// (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).
//
// In host code `scratch` names a host-side stand-in, not a device address;
// its device address differs per GPU and is resolved once per device with
// hipGetSymbolAddress (the device current at that moment) and cached.  The
// library resolves it when a context is created (amfls_context_create):
// the first resolution loads this file's device code and its array, which
// must not happen inside the allocation-free entry point.  All
// calls on one GPU share its array: the kernels that use it write every
// entry they later read (packed block planes, per-chunk partial sums), so
// nothing needs clearing, and calls that overlap on one GPU race.
#include "kernels.hpp"

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

namespace amflsx {
namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void*, kMaxDevices> g_ptr{};

}  // namespace

Scratch noalloc_scratch() {
    int dev = 0;
    AMFLS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range: " + std::to_string(dev));
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (g_ptr[dev] == nullptr) {
            void* addr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || addr == nullptr)
                throw std::runtime_error(std::string("hipGetSymbolAddress of the allocation-free scratch failed: ") +
                                         hipGetErrorString(e));
            g_ptr[dev] = addr;
        }
        p = g_ptr[dev];
    }
    Scratch s;
    s.ptr = p;
    s.bytes = kNoAllocScratchBytes;
    s.fixed = true;
    return s;
}

}  // namespace amflsx
