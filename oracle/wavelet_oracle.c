/*
 * CPU oracle for the wavelet BayesShrink chroma denoise node
 * (com.cpipe.denoise.wavelet_bayes_shrink) -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/ and bench.py's cpu_baseline leg may load it, as the reference result.
 *
 * Restates the reference Halide generator
 * src/cpipe/nodes/denoise/wavelet_bayes_shrink_generator.cpp:9-31,57-86 in
 * FP32 with its operation order and no fused multiply-add (the reference's
 * x86-64 target has no FMA):
 *   repeat_edge input (:61); luma = ((r + 2g) + b) * 0.25, co = r - b,
 *   cg = g - (r + b) * 0.5 (:63-71);
 *   per 2x2 block (bx = 2 floor(x/2), :17-22) of a chroma plane:
 *   ll = 0.25 (((p00 + p10) + p01) + p11) and the three detail bands
 *   h = 0.25 (((p00 - p10) + p01) - p11), v = 0.25 (((p00 + p10) - p01) - p11),
 *   d = 0.25 (((p00 - p10) - p01) + p11), each soft-thresholded at
 *   0.035 * chroma_strength (:9-12,24-27,75); the pixel's value is the inverse
 *   Haar of its position in the block (:29-30);
 *   t = luma - 0.5 dcg; r = t + 0.5 dco, g = t + dcg, b = t - 0.5 dco, alpha
 *   passes through (:78-83); every channel max(0, .) (taken as v > 0 ? v : +0)
 *   and cast to binary16, round to nearest even (:84).
 * Parity pins: tests/test_oracle_wavelet.py (the reference unit test
 * tests/unit/test_node_denoise_wavelet_bayes_shrink.cpp:38-87, constant images
 * reproduced bit for bit, strength 0 as the identity, the repeat-edge rule,
 * full attenuation of sub-threshold detail).
 */
#include <stddef.h>
#include <stdint.h>

#include "bm3d_oracle.h"

static float soft_threshold(float value, float threshold) {
    if (value > threshold) return value - threshold;
    if (value < -threshold) return value + threshold;
    return 0.0f;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

typedef struct {
    const uint16_t* in;
    size_t stride;
    int w, h;
} rgba_view;

static float sample(const rgba_view* v, int x, int y, int c) {
    x = clampi(x, 0, v->w - 1);
    y = clampi(y, 0, v->h - 1);
    return oracle_f16_to_f32(v->in[((size_t)y * v->stride + (size_t)x) * 4 + (size_t)c]);
}

/* component 0: co = r - b; 1: cg = g - (r + b) * 0.5 */
static float chroma(const rgba_view* v, int x, int y, int component) {
    const float r = sample(v, x, y, 0), g = sample(v, x, y, 1), b = sample(v, x, y, 2);
    return component == 0 ? r - b : g - ((r + b) * 0.5f);
}

static float block_value(const rgba_view* v, int x, int y, int component, float threshold) {
    const int bx = (x / 2) * 2, by = (y / 2) * 2;
    const float p00 = chroma(v, bx, by, component), p10 = chroma(v, bx + 1, by, component);
    const float p01 = chroma(v, bx, by + 1, component), p11 = chroma(v, bx + 1, by + 1, component);
    const float ll = 0.25f * (((p00 + p10) + p01) + p11);
    const float hb = soft_threshold(0.25f * (((p00 - p10) + p01) - p11), threshold);
    const float vb = soft_threshold(0.25f * (((p00 + p10) - p01) - p11), threshold);
    const float db = soft_threshold(0.25f * (((p00 - p10) - p01) + p11), threshold);
    const int ox = x & 1, oy = y & 1;
    if (ox == 0 && oy == 0) return ((ll + hb) + vb) + db;
    if (ox == 1 && oy == 0) return ((ll - hb) + vb) - db;
    if (ox == 0 && oy == 1) return ((ll + hb) - vb) - db;
    return ((ll - hb) - vb) + db;
}

static uint16_t out_half(float v) { return oracle_f32_to_f16(v > 0.0f ? v : 0.0f); }

void oracle_wavelet_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                            size_t out_stride_px, int width, int height, float chroma_strength) {
    const rgba_view view = {in, in_stride_px, width, height};
    const float threshold = 0.035f * chroma_strength;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            const float r = sample(&view, x, y, 0), g = sample(&view, x, y, 1), b = sample(&view, x, y, 2);
            const float luma = ((r + (2.0f * g)) + b) * 0.25f;
            const float dco = block_value(&view, x, y, 0, threshold);
            const float dcg = block_value(&view, x, y, 1, threshold);
            const float t = luma - (0.5f * dcg);
            uint16_t* o = out + ((size_t)y * out_stride_px + (size_t)x) * 4;
            o[0] = out_half(t + (0.5f * dco));
            o[1] = out_half(t + dcg);
            o[2] = out_half(t - (0.5f * dco));
            o[3] = out_half(sample(&view, x, y, 3));
        }
    }
}
