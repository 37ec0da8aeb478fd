/*
 * cpipe_convert.h -- C ABI of the MI355X (gfx950) precision conversions of the
 * precision-convert node (com.cpipe.precision_convert), the implicit node the
 * reference's precision planner inserts between nodes whose formats differ.
 *
 * Plain C: pointers, sizes and status codes (cpipe_status_t: 0 OK, 3 OOM,
 * 4 BAD_PRECISION, 5 BAD_INDEX, 7 INTERNAL_ERROR, 8 UNSUPPORTED).
 *
 * What each entry point replaces in the reference (xndcn/cpipe):
 *   - cpipe_convert: the Halide AOT filters precision_convert_r16u_to_f32,
 *     precision_convert_f32_to_rgba16f, precision_convert_rgba16f_to_rgba8 and
 *     precision_convert_rgba16f_to_rgba16u
 *     (src/cpipe/nodes/precision/precision_convert_generator.cpp:31-114), chosen by
 *     filter_for (src/cpipe/nodes/precision/precision_convert.cpp:13-28);
 *   - cpipe_convert_host: the same on host buffers, for the node;
 *   - the node itself replaces PrecisionConvert (precision_convert.cpp:40-80) and its
 *     CPIPE_REGISTER_NODE (:86); its descriptor sits in the engine library's
 *     `cpipe_registry` section.
 */
#ifndef CPIPE_CONVERT_H
#define CPIPE_CONVERT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pixel formats, numbered as the reference's PixelFormat
 * (include/cpipe/core/PixelFormat.hpp:11-29). */
enum {
    CPIPE_CONVERT_R16_UINT = 1,
    CPIPE_CONVERT_R8G8B8A8_UNORM = 3,
    CPIPE_CONVERT_R16G16B16A16_SFLOAT = 6,
    CPIPE_CONVERT_R32_SFLOAT = 10,
    CPIPE_CONVERT_R16G16B16A16_UNORM = 16
};

/* 1 when in_format -> out_format is one of the four conversions, else 0. */
int cpipe_convert_supported(int in_format, int out_format);

/* Device pointers on the current HIP device, row strides in bytes; buffers and
 * strides aligned to the pixel size (2, 4 or 8 bytes). R16_UINT -> R32_SFLOAT
 * (value as float); R32_SFLOAT -> R16G16B16A16_SFLOAT (grey to RGB, alpha 1);
 * R16G16B16A16_SFLOAT -> R8G8B8A8_UNORM / R16G16B16A16_UNORM
 * ((uint)(clamp(v, 0, 1) * max + 0.5)). Enqueues on `stream` (hipStream_t as
 * void*, NULL = default). An unsupported pair gives 4 (BAD_PRECISION). */
int cpipe_convert(int in_format, const void* in, size_t in_stride_bytes, int out_format, void* out,
                  size_t out_stride_bytes, int width, int height, void* stream);

/* The same on host memory (HIP device 0): copies in, runs, copies out;
 * synchronous. Device and pinned host staging buffers are kept between calls
 * (grow-only, process-wide); rows are copied into the pinned staging on up to
 * 8 threads, so PCIe runs at DMA rate. */
int cpipe_convert_host(int in_format, const void* in, size_t in_stride_bytes, int out_format,
                       void* out, size_t out_stride_bytes, int width, int height);

/* Last error message of the calling thread (never NULL). */
const char* cpipe_convert_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
