// Wavelet BayesShrink chroma denoise (com.cpipe.denoise.wavelet_bayes_shrink)
// for gfx950, and the C ABI of include/cpipe_wavelet.h.
//
// The arithmetic of the reference Halide generator
// src/cpipe/nodes/denoise/wavelet_bayes_shrink_generator.cpp:9-31,57-86, in
// FP32 with its operation order and no contraction (-ffp-contract=off; the
// reference's x86-64 target has no FMA):
//   luma = ((r + 2g) + b) * 0.25, co = r - b, cg = g - (r + b) * 0.5 (:63-71);
//   per 2x2 block of each chroma plane: ll = 0.25 (((p00 + p10) + p01) + p11),
//   h = 0.25 (((p00 - p10) + p01) - p11), v = 0.25 (((p00 + p10) - p01) - p11),
//   d = 0.25 (((p00 - p10) - p01) + p11), h/v/d soft-thresholded at
//   0.035 * chroma_strength (:9-12,24-27,75), the pixel rebuilt by the inverse
//   Haar of its block position (:29-30);
//   t = luma - 0.5 dcg, r = t + 0.5 dco, g = t + dcg, b = t - 0.5 dco, alpha
//   kept (:78-83); max(0, .) as v > 0 ? v : +0, binary16 round to nearest
//   even (:84); reads past an odd edge repeat the last row / column (:61).
//
// One thread per 2x2 block: it loads the block's four pixels once (8 bytes
// each; a block's two pixels of a row are adjacent, so a wave reads and writes
// contiguous 1 KiB row spans), computes both chroma planes' bands and writes
// the four output pixels. HBM-bound: 16 bytes per pixel (8 in, 8 out).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cpipe_wavelet.h"
#include "host_staging.h"

namespace {

enum Status { OK = 0, OOM = 3, BAD_INDEX = 5, INTERNAL = 7, UNSUPPORTED = 8 };

thread_local std::string g_werr;

int fail(int code, const std::string& msg) {
    g_werr = msg;
    return code;
}

#define WHIP_CHECK(expr)                                                                \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? OOM : INTERNAL,                     \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

constexpr int WX = 64, WY = 4;  // blocks per workgroup (128 x 8 pixels)

struct Px {
    float c[4];
};

__device__ __forceinline__ Px load_px(const uint2* __restrict__ p) {
    const uint2 v = *p;
    Px o;
    o.c[0] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.x & 0xffffu));
    o.c[1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.x >> 16));
    o.c[2] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y & 0xffffu));
    o.c[3] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y >> 16));
    return o;
}

__device__ __forceinline__ uint32_t half_bits(float v) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(v > 0.0f ? v : 0.0f));
}

__device__ __forceinline__ float soft(float value, float threshold) {
    return value > threshold ? value - threshold : (value < -threshold ? value + threshold : 0.0f);
}

// Thresholded bands of one chroma plane over a block: ll, h, v, d.
__device__ __forceinline__ void bands(float p00, float p10, float p01, float p11, float thr, float& ll,
                                      float& h, float& v, float& d) {
    ll = 0.25f * (((p00 + p10) + p01) + p11);
    h = soft(0.25f * (((p00 - p10) + p01) - p11), thr);
    v = soft(0.25f * (((p00 + p10) - p01) - p11), thr);
    d = soft(0.25f * (((p00 - p10) - p01) + p11), thr);
}

__device__ __forceinline__ uint2 out_px(const Px& p, float dco, float dcg) {
    const float luma = ((p.c[0] + (2.0f * p.c[1])) + p.c[2]) * 0.25f;
    const float t = luma - (0.5f * dcg);
    const float r = t + (0.5f * dco), g = t + dcg, b = t - (0.5f * dco);
    return make_uint2(half_bits(r) | (half_bits(g) << 16), half_bits(b) | (half_bits(p.c[3]) << 16));
}

__global__ __launch_bounds__(WX * WY) void k_wavelet(const uint16_t* __restrict__ in, size_t in_stride,
                                                     uint16_t* __restrict__ out, size_t out_stride,
                                                     int w, int h, float thr) {
    const int bxi = blockIdx.x * WX + (int)(threadIdx.x % WX);
    const int byi = blockIdx.y * WY + (int)(threadIdx.x / WX);
    const int x0 = 2 * bxi, y0 = 2 * byi;
    if (x0 >= w || y0 >= h) return;
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);  // repeat_edge
    const uint2* src = reinterpret_cast<const uint2*>(in);
    const Px q00 = load_px(src + (size_t)y0 * in_stride + x0);
    const Px q10 = load_px(src + (size_t)y0 * in_stride + x1);
    const Px q01 = load_px(src + (size_t)y1 * in_stride + x0);
    const Px q11 = load_px(src + (size_t)y1 * in_stride + x1);
    auto co = [](const Px& p) { return p.c[0] - p.c[2]; };
    auto cg = [](const Px& p) { return p.c[1] - ((p.c[0] + p.c[2]) * 0.5f); };
    float oll, oh, ov, od, gll, gh, gv, gd;
    bands(co(q00), co(q10), co(q01), co(q11), thr, oll, oh, ov, od);
    bands(cg(q00), cg(q10), cg(q01), cg(q11), thr, gll, gh, gv, gd);
    uint2* dst = reinterpret_cast<uint2*>(out);
    dst[(size_t)y0 * out_stride + x0] = out_px(q00, ((oll + oh) + ov) + od, ((gll + gh) + gv) + gd);
    if (x0 + 1 < w)
        dst[(size_t)y0 * out_stride + x0 + 1] = out_px(q10, ((oll - oh) + ov) - od, ((gll - gh) + gv) - gd);
    if (y0 + 1 < h) {
        dst[(size_t)(y0 + 1) * out_stride + x0] = out_px(q01, ((oll + oh) - ov) - od, ((gll + gh) - gv) - gd);
        if (x0 + 1 < w)
            dst[(size_t)(y0 + 1) * out_stride + x0 + 1] =
                out_px(q11, ((oll - oh) - ov) + od, ((gll - gh) - gv) + gd);
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

extern "C" int cpipe_wavelet_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                                     size_t out_stride_px, int width, int height, float chroma_strength,
                                     void* stream) {
    if (width <= 0 || height <= 0) return fail(BAD_INDEX, "wavelet: empty image");
    if (in == nullptr || out == nullptr) return fail(BAD_INDEX, "wavelet: null buffer");
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 7u)
        return fail(BAD_INDEX, "wavelet: buffers must be 8-byte aligned (one pixel)");
    if (in_stride_px < (size_t)width || out_stride_px < (size_t)width)
        return fail(BAD_INDEX, "wavelet: row stride below the width");
    if (!(chroma_strength >= 0.0f)) return fail(BAD_INDEX, "wavelet: chroma_strength must be >= 0");
    const dim3 grid((unsigned)(((width + 1) / 2 + WX - 1) / WX), (unsigned)(((height + 1) / 2 + WY - 1) / WY));
    k_wavelet<<<grid, WX * WY, 0, static_cast<hipStream_t>(stream)>>>(in, in_stride_px, out, out_stride_px,
                                                                     width, height, 0.035f * chroma_strength);
    WHIP_CHECK(hipGetLastError());
    return OK;
}

extern "C" int cpipe_wavelet_host_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                                          size_t out_stride_px, int width, int height,
                                          float chroma_strength) {
    if (width <= 0 || height <= 0) return fail(BAD_INDEX, "wavelet: empty image");
    if (in == nullptr || out == nullptr) return fail(BAD_INDEX, "wavelet: null buffer");
    if (in_stride_px < (size_t)width || out_stride_px < (size_t)width)
        return fail(BAD_INDEX, "wavelet: row stride below the width");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(UNSUPPORTED, "wavelet: no HIP device visible");
    // host buffers run on HIP device 0 (the node's device); the caller's
    // current device is restored on return
    cpipe_internal::ScopedDevice device(0);
    if (!device.ok()) return fail(UNSUPPORTED, "wavelet: cannot select HIP device 0");
    HostBuffers& hb = host_buffers();
    std::lock_guard<std::mutex> lock(hb.mu);
    hipStream_t st = nullptr;
    WHIP_CHECK(hb.stream.get(&st));
    const size_t row = (size_t)width * 8, bytes = row * (size_t)height;
    WHIP_CHECK(hb.in.reserve(bytes));
    WHIP_CHECK(hb.out.reserve(bytes));
    cpipe_internal::par_rows(hb.in.pin, row, in, in_stride_px * 8, row, height);
    WHIP_CHECK(hipMemcpyAsync(hb.in.dev, hb.in.pin, bytes, hipMemcpyHostToDevice, st));
    int rc = cpipe_wavelet_rgba16f(static_cast<const uint16_t*>(hb.in.dev), (size_t)width,
                                   static_cast<uint16_t*>(hb.out.dev), (size_t)width, width, height,
                                   chroma_strength, st);
    if (rc != OK) return rc;
    WHIP_CHECK(hipMemcpyAsync(hb.out.pin, hb.out.dev, bytes, hipMemcpyDeviceToHost, st));
    WHIP_CHECK(hipStreamSynchronize(st));
    cpipe_internal::par_rows(out, out_stride_px * 8, hb.out.pin, row, row, height);
    return OK;
}

extern "C" const char* cpipe_wavelet_last_error(void) { return g_werr.c_str(); }
