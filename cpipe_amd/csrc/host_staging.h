// Host-memory staging shared by the host-buffer entry points of the engine
// (cpipe_bm3d_host_*, cpipe_wavelet_host_rgba16f, cpipe_convert_host): the
// caller's buffers are pageable, so rows are copied into pinned staging on up to
// 8 threads and the DMA engines then move whole images at full PCIe rate.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace cpipe_internal {

// Row copy between host buffers, split over up to 8 threads above 8 MiB.
inline void par_rows(void* dst, size_t dpitch, const void* src, size_t spitch, size_t row_bytes, int rows) {
    auto band = [=](int r0, int r1) {
        if (dpitch == row_bytes && spitch == row_bytes) {
            std::memcpy(static_cast<char*>(dst) + (size_t)r0 * dpitch,
                        static_cast<const char*>(src) + (size_t)r0 * spitch, (size_t)(r1 - r0) * row_bytes);
            return;
        }
        for (int r = r0; r < r1; ++r)
            std::memcpy(static_cast<char*>(dst) + (size_t)r * dpitch,
                        static_cast<const char*>(src) + (size_t)r * spitch, row_bytes);
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = (size_t)rows * row_bytes < ((size_t)8 << 20) ? 1 : (int)std::min(8u, hw ? hw : 1u);
    if (nt <= 1) {
        band(0, rows);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t)
        th.emplace_back(band, (int)((int64_t)rows * t / nt), (int)((int64_t)rows * (t + 1) / nt));
    band(0, rows / nt);
    for (auto& x : th) x.join();
}

// Makes `dev` the current HIP device for a scope and restores the caller's
// current device on every return path.
class ScopedDevice {
public:
    explicit ScopedDevice(int dev) {
        if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
        ok_ = prev_ == dev || hipSetDevice(dev) == hipSuccess;
    }
    ~ScopedDevice() {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }
    bool ok() const { return ok_; }

private:
    int prev_ = -1;
    bool ok_ = false;
};

// A non-blocking stream owned by a host-buffer entry point (created on first
// use on the current device, kept until the process ends): its copies and
// kernels are ordered among themselves only, so a host-buffer call neither
// waits for nor holds up work that other callers queued on the device.
struct OwnStream {
    hipStream_t st = nullptr;
    hipError_t get(hipStream_t* out) {
        if (!st) {
            const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e != hipSuccess) {
                st = nullptr;
                return e;
            }
        }
        *out = st;
        return hipSuccess;
    }
};

// A grow-only pair of device and pinned host buffers of equal size (kept
// until the process ends; the caller serialises use).
struct StagedBuffer {
    void* dev = nullptr;
    void* pin = nullptr;
    size_t bytes = 0;

    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (dev) (void)hipFree(dev);
        if (pin) (void)hipHostFree(pin);
        dev = pin = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&dev, n);
        if (e == hipSuccess) e = hipHostMalloc(&pin, n, hipHostMallocDefault);
        if (e == hipSuccess) bytes = n;
        return e;
    }
};

}  // namespace cpipe_internal
