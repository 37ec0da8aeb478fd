// Precision conversions of the precision-convert node (com.cpipe.precision_convert)
// for gfx950, and the C ABI of include/cpipe_convert.h.
//
// The four reference Halide generators of
// src/cpipe/nodes/precision/precision_convert_generator.cpp, one kernel each
// (a thread per pixel, rows coalesced):
//   R16_UINT -> R32_SFLOAT             out = (float)u16                          (:38)
//   R32_SFLOAT -> R16G16B16A16_SFLOAT  out = (f16 v, f16 v, f16 v, f16 1)         (:54-55)
//   R16G16B16A16_SFLOAT -> R8G8B8A8_UNORM      (uint8)(clamp(v, 0, 1) * 255 + 0.5)   (:12-14,74-76)
//   R16G16B16A16_SFLOAT -> R16G16B16A16_UNORM  (uint16)(clamp(v, 0, 1) * 65535 + 0.5) (:12-14,95-97)
// FP32 multiply then add (no contraction), float -> integer casts truncating
// toward zero, binary16 round to nearest even; clamp as min(max(v, 0), 1) with
// NaN -> 0. All four are HBM-bound copies.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cpipe_convert.h"
#include "host_staging.h"

namespace {

enum Status { OK = 0, OOM = 3, BAD_PRECISION = 4, BAD_INDEX = 5, INTERNAL = 7, UNSUPPORTED = 8 };

thread_local std::string g_cerr;

int fail(int code, const std::string& msg) {
    g_cerr = msg;
    return code;
}

#define CHIP_CHECK(expr)                                                                \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? OOM : INTERNAL,                     \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

constexpr int CT = 256;  // threads per workgroup: one row segment of 256 pixels

__device__ __forceinline__ float half_to_float(uint32_t b) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xffffu));
}

__device__ __forceinline__ float clamp_unit(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }

__global__ __launch_bounds__(CT) void k_r16u_to_f32(const uint8_t* __restrict__ in, size_t is,
                                                    uint8_t* __restrict__ out, size_t os, int w) {
    const int x = blockIdx.x * CT + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint16_t v = reinterpret_cast<const uint16_t*>(in + (size_t)y * is)[x];
    reinterpret_cast<float*>(out + (size_t)y * os)[x] = (float)v;
}

__global__ __launch_bounds__(CT) void k_f32_to_rgba16f(const uint8_t* __restrict__ in, size_t is,
                                                       uint8_t* __restrict__ out, size_t os, int w) {
    const int x = blockIdx.x * CT + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const float v = reinterpret_cast<const float*>(in + (size_t)y * is)[x];
    const uint32_t hv = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
    const uint32_t one = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)1.0f);
    reinterpret_cast<uint2*>(out + (size_t)y * os)[x] = make_uint2(hv | (hv << 16), hv | (one << 16));
}

template <int BITS>
__global__ __launch_bounds__(CT) void k_rgba16f_to_unorm(const uint8_t* __restrict__ in, size_t is,
                                                         uint8_t* __restrict__ out, size_t os, int w) {
    const int x = blockIdx.x * CT + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint2 p = reinterpret_cast<const uint2*>(in + (size_t)y * is)[x];
    const float c[4] = {half_to_float(p.x), half_to_float(p.x >> 16), half_to_float(p.y),
                        half_to_float(p.y >> 16)};
    constexpr float scale = BITS == 8 ? 255.0f : 65535.0f;
    uint32_t q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = (uint32_t)((clamp_unit(c[k]) * scale) + 0.5f);
    if (BITS == 8)
        reinterpret_cast<uint32_t*>(out + (size_t)y * os)[x] = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
    else
        reinterpret_cast<uint2*>(out + (size_t)y * os)[x] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
}

// Bytes per pixel of the formats the node handles (0: unsupported).
int bytes_per_pixel(int format) {
    switch (format) {
        case CPIPE_CONVERT_R16_UINT: return 2;
        case CPIPE_CONVERT_R8G8B8A8_UNORM: return 4;
        case CPIPE_CONVERT_R16G16B16A16_SFLOAT: return 8;
        case CPIPE_CONVERT_R32_SFLOAT: return 4;
        case CPIPE_CONVERT_R16G16B16A16_UNORM: return 8;
        default: return 0;
    }
}

// Device and pinned staging of the host-memory entry point (grow-only, process-wide).
struct HostBuffers {
    std::mutex mu;
    cpipe_internal::StagedBuffer in, out;
    cpipe_internal::OwnStream stream;
};

HostBuffers& host_buffers() {
    static HostBuffers* b = new HostBuffers();  // kept until the process ends
    return *b;
}

}  // namespace

extern "C" int cpipe_convert_supported(int in_format, int out_format) {
    return (in_format == CPIPE_CONVERT_R16_UINT && out_format == CPIPE_CONVERT_R32_SFLOAT) ||
           (in_format == CPIPE_CONVERT_R32_SFLOAT && out_format == CPIPE_CONVERT_R16G16B16A16_SFLOAT) ||
           (in_format == CPIPE_CONVERT_R16G16B16A16_SFLOAT &&
            (out_format == CPIPE_CONVERT_R8G8B8A8_UNORM || out_format == CPIPE_CONVERT_R16G16B16A16_UNORM));
}

extern "C" int cpipe_convert(int in_format, const void* in, size_t in_stride_bytes, int out_format, void* out,
                             size_t out_stride_bytes, int width, int height, void* stream) {
    if (!cpipe_convert_supported(in_format, out_format))
        return fail(BAD_PRECISION, "convert: unsupported format pair");
    if (width <= 0 || height <= 0) return fail(BAD_INDEX, "convert: empty image");
    if (in == nullptr || out == nullptr) return fail(BAD_INDEX, "convert: null buffer");
    const int ib = bytes_per_pixel(in_format), ob = bytes_per_pixel(out_format);
    const size_t ia = ib == 2 ? 2 : (ib == 4 ? 4 : 8), oa = ob == 2 ? 2 : (ob == 4 ? 4 : 8);
    if ((reinterpret_cast<uintptr_t>(in) % ia) || (reinterpret_cast<uintptr_t>(out) % oa) || in_stride_bytes % ia ||
        out_stride_bytes % oa)
        return fail(BAD_INDEX, "convert: buffers and strides must be aligned to the pixel size");
    if (in_stride_bytes < (size_t)width * ib || out_stride_bytes < (size_t)width * ob)
        return fail(BAD_INDEX, "convert: row stride below the width");
    const dim3 grid((unsigned)((width + CT - 1) / CT), (unsigned)height);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const uint8_t* s = static_cast<const uint8_t*>(in);
    uint8_t* d = static_cast<uint8_t*>(out);
    if (in_format == CPIPE_CONVERT_R16_UINT)
        k_r16u_to_f32<<<grid, CT, 0, st>>>(s, in_stride_bytes, d, out_stride_bytes, width);
    else if (in_format == CPIPE_CONVERT_R32_SFLOAT)
        k_f32_to_rgba16f<<<grid, CT, 0, st>>>(s, in_stride_bytes, d, out_stride_bytes, width);
    else if (out_format == CPIPE_CONVERT_R8G8B8A8_UNORM)
        k_rgba16f_to_unorm<8><<<grid, CT, 0, st>>>(s, in_stride_bytes, d, out_stride_bytes, width);
    else
        k_rgba16f_to_unorm<16><<<grid, CT, 0, st>>>(s, in_stride_bytes, d, out_stride_bytes, width);
    CHIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int cpipe_convert_host(int in_format, const void* in, size_t in_stride_bytes, int out_format,
                                  void* out, size_t out_stride_bytes, int width, int height) {
    if (!cpipe_convert_supported(in_format, out_format))
        return fail(BAD_PRECISION, "convert: unsupported format pair");
    if (width <= 0 || height <= 0) return fail(BAD_INDEX, "convert: empty image");
    if (in == nullptr || out == nullptr) return fail(BAD_INDEX, "convert: null buffer");
    const size_t irow = (size_t)width * bytes_per_pixel(in_format), orow = (size_t)width * bytes_per_pixel(out_format);
    if (in_stride_bytes < irow || out_stride_bytes < orow) return fail(BAD_INDEX, "convert: row stride below the width");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(UNSUPPORTED, "convert: no HIP device visible");
    // host buffers run on HIP device 0 (the node's device); the caller's
    // current device is restored on return
    cpipe_internal::ScopedDevice device(0);
    if (!device.ok()) return fail(UNSUPPORTED, "convert: cannot select HIP device 0");
    HostBuffers& hb = host_buffers();
    std::lock_guard<std::mutex> lock(hb.mu);
    hipStream_t st = nullptr;
    CHIP_CHECK(hb.stream.get(&st));
    const size_t ib = irow * (size_t)height, ob = orow * (size_t)height;
    CHIP_CHECK(hb.in.reserve(ib));
    CHIP_CHECK(hb.out.reserve(ob));
    cpipe_internal::par_rows(hb.in.pin, irow, in, in_stride_bytes, irow, height);
    CHIP_CHECK(hipMemcpyAsync(hb.in.dev, hb.in.pin, ib, hipMemcpyHostToDevice, st));
    const int rc = cpipe_convert(in_format, hb.in.dev, irow, out_format, hb.out.dev, orow, width, height, st);
    if (rc != OK) return rc;
    CHIP_CHECK(hipMemcpyAsync(hb.out.pin, hb.out.dev, ob, hipMemcpyDeviceToHost, st));
    CHIP_CHECK(hipStreamSynchronize(st));
    cpipe_internal::par_rows(out, out_stride_bytes, hb.out.pin, orow, orow, height);
    return OK;
}

extern "C" const char* cpipe_convert_last_error(void) { return g_cerr.c_str(); }
