/*
 * cpipe_wavelet.h -- C ABI of the MI355X (gfx950) wavelet BayesShrink chroma
 * denoise, the node downstream of BM3D in the reference's classic pipeline
 * (examples/pipelines/full-classic-pipeline.cpipe.json:127-129).
 *
 * Plain C: pointers, sizes and status codes (cpipe_status_t: 0 OK, 3 OOM,
 * 5 BAD_INDEX, 7 INTERNAL_ERROR, 8 UNSUPPORTED); no HIP or torch types.
 *
 * What each entry point replaces in the reference (xndcn/cpipe):
 *   - cpipe_wavelet_rgba16f: the Halide AOT filter `denoise_wavelet_bayes_shrink`
 *     (src/cpipe/nodes/denoise/wavelet_bayes_shrink_generator.cpp:57-86), reached
 *     through its param-blob adapter denoise_wavelet_bayes_shrink_param
 *     (src/cpipe/nodes/denoise/wavelet_bayes_shrink_dispatch.cpp:19-28) with the
 *     4-byte blob {float chroma_strength} (:12-14).
 *   - cpipe_wavelet_host_rgba16f: the same on host buffers, for the node
 *     (the IBuffer lock_cpu contract of HalideBufferAdapter.cpp:41).
 *   - the node itself, com.cpipe.denoise.wavelet_bayes_shrink, replaces
 *     DenoiseWaveletBayesShrink (src/cpipe/nodes/denoise/wavelet_bayes_shrink.cpp:27-47)
 *     and its CPIPE_REGISTER_NODE (:54-55); its descriptor sits in the engine
 *     library's `cpipe_registry` section next to the BM3D node's.
 */
#ifndef CPIPE_WAVELET_H
#define CPIPE_WAVELET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interleaved RGBA binary16 in and out (8 bytes per pixel), device pointers on
 * the current HIP device, row strides in pixels (>= width). Soft threshold
 * 0.035 * chroma_strength on the three Haar detail bands of Co = R - B and
 * Cg = G - (R + B) / 2 per 2x2 block, edges repeated, luma and alpha kept,
 * every channel clamped at 0. `stream` is a hipStream_t passed as void*
 * (NULL = the default stream); the call only enqueues. in and out must be
 * 8-byte aligned and must not overlap. */
int cpipe_wavelet_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                          size_t out_stride_px, int width, int height, float chroma_strength,
                          void* stream);

/* The same on host memory (HIP device 0): copies in, runs, copies out;
 * synchronous. Device and pinned host staging buffers are kept between calls
 * (grow-only, process-wide); rows are copied into the pinned staging on up to
 * 8 threads, so PCIe runs at DMA rate. */
int cpipe_wavelet_host_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                               size_t out_stride_px, int width, int height,
                               float chroma_strength);

/* Last error message of the calling thread (never NULL). */
const char* cpipe_wavelet_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
