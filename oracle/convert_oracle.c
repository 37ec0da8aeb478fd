/*
 * CPU oracle for the precision-convert node (com.cpipe.precision_convert) --
 * TEST INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline leg).
 *
 * Restates the reference Halide generators
 * src/cpipe/nodes/precision/precision_convert_generator.cpp:
 *   r16u_to_f32      out = (float)u16                                   (:38)
 *   f32_to_rgba16f   out = (f16(v), f16(v), f16(v), f16(1))             (:54-55)
 *   rgba16f_to_rgba8 out = (uint8)(clamp(v, 0, 1) * 255 + 0.5)           (:12-14,74-76)
 *   rgba16f_to_rgba16u out = (uint16)(clamp(v, 0, 1) * 65535 + 0.5)     (:12-14,95-97)
 * in FP32, multiply then add (no fused multiply-add on the reference's x86-64
 * target), the integer cast truncating toward zero (Halide's float -> uint
 * cast); binary16 round to nearest even. clamp is min(max(v, 0), 1) with NaN
 * taken as 0 (the reference's NaN behaviour is its codegen's; unpinned).
 * Strides in bytes. Parity pins: tests/test_oracle_convert.py.
 */
#include <stddef.h>
#include <stdint.h>

#include "bm3d_oracle.h"

static float clamp_unit(float v) {
    if (!(v > 0.0f)) return 0.0f; /* also NaN */
    return v < 1.0f ? v : 1.0f;
}

void oracle_convert_r16u_to_f32(const uint8_t* in, size_t in_stride, uint8_t* out, size_t out_stride,
                                int width, int height) {
    for (int y = 0; y < height; ++y) {
        const uint16_t* s = (const uint16_t*)(in + (size_t)y * in_stride);
        float* d = (float*)(out + (size_t)y * out_stride);
        for (int x = 0; x < width; ++x) d[x] = (float)s[x];
    }
}

void oracle_convert_f32_to_rgba16f(const uint8_t* in, size_t in_stride, uint8_t* out,
                                   size_t out_stride, int width, int height) {
    const uint16_t one = oracle_f32_to_f16(1.0f);
    for (int y = 0; y < height; ++y) {
        const float* s = (const float*)(in + (size_t)y * in_stride);
        uint16_t* d = (uint16_t*)(out + (size_t)y * out_stride);
        for (int x = 0; x < width; ++x) {
            const uint16_t v = oracle_f32_to_f16(s[x]);
            d[4 * x + 0] = v;
            d[4 * x + 1] = v;
            d[4 * x + 2] = v;
            d[4 * x + 3] = one;
        }
    }
}

void oracle_convert_rgba16f_to_unorm(const uint8_t* in, size_t in_stride, uint8_t* out,
                                     size_t out_stride, int width, int height, int bits) {
    const float scale = bits == 8 ? 255.0f : 65535.0f;
    for (int y = 0; y < height; ++y) {
        const uint16_t* s = (const uint16_t*)(in + (size_t)y * in_stride);
        for (int x = 0; x < 4 * width; ++x) {
            const float scaled = (clamp_unit(oracle_f16_to_f32(s[x])) * scale) + 0.5f;
            if (bits == 8)
                out[(size_t)y * out_stride + (size_t)x] = (uint8_t)scaled;
            else
                ((uint16_t*)(out + (size_t)y * out_stride))[x] = (uint16_t)scaled;
        }
    }
}
