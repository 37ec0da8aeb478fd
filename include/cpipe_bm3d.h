/*
 * cpipe_bm3d.h -- C ABI of the MI355X (gfx950) BM3D denoise engine.
 *
 * Plain C: pointers, sizes and status codes; no HIP or torch types. Device
 * pointers are HIP device addresses on the plan's device; `stream` is a
 * hipStream_t passed as void* (NULL = the default stream). Status values are
 * cpipe_status_t codes (include/cpipe_node_abi.h): 0 OK, 3 OOM, 5 BAD_INDEX,
 * 7 INTERNAL_ERROR, 8 UNSUPPORTED.
 *
 * What each entry point replaces in the reference (xndcn/cpipe):
 *   - cpipe_bm3d_compat_rgba16f: the Halide AOT filters `denoise_bm3d_step1`
 *     (src/cpipe/nodes/denoise/bm3d_step1_generator.cpp:42-56) and
 *     `denoise_bm3d_step2` (bm3d_step2_generator.cpp:33-42), reached through the
 *     param-blob adapters denoise_bm3d_step{1,2}_param
 *     (src/cpipe/nodes/denoise/bm3d_dispatch.cpp:19-41).
 *   - cpipe_bm3d_plan_* : the two-step BM3D the reference only describes
 *     (docs/research/07-classic-isp-algorithms.md:277-283,541-560) and that its
 *     node submits through ComputeContext::submit_halide_with_params
 *     (src/cpipe/nodes/denoise/bm3d.cpp:96-100). It reads/writes the FP32
 *     planes of an IBuffer (include/cpipe/core/PixelFormat.hpp:22, R32_SFLOAT).
 *   - cpipe_bm3d_node_descriptor: the CPIPE_REGISTER_NODE descriptor of
 *     src/cpipe/nodes/denoise/bm3d.cpp:111 (include/cpipe/sdk/registry.hpp:14-32).
 */
#ifndef CPIPE_BM3D_H
#define CPIPE_BM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPIPE_BM3D_API_VERSION 1

/* Arithmetic of the block-matching cross term and group transforms. */
enum {
    CPIPE_BM3D_PRECISION_FP32 = 0, /* exact FP32 (fma order fixed, matches the CPU FP32 path) */
    CPIPE_BM3D_PRECISION_BF16 = 1, /* BF16 MFMA cross term + FP32 re-rank */
    CPIPE_BM3D_PRECISION_FP16 = 2  /* FP16 MFMA cross term + FP32 re-rank */
};

/* Stage timers reported by cpipe_bm3d_plan_timing (milliseconds, summed over
 * all runs since timing was enabled). */
enum {
    CPIPE_BM3D_T_COLOR = 0,   /* opponent transform in */
    CPIPE_BM3D_T_BM1 = 1,     /* block matching, step 1 */
    CPIPE_BM3D_T_GROUP1 = 2,  /* transform + hard threshold + tile aggregation */
    CPIPE_BM3D_T_MERGE1 = 3,  /* tile merge + normalisation -> basic estimate */
    CPIPE_BM3D_T_BM2 = 4,     /* block matching, step 2 */
    CPIPE_BM3D_T_GROUP2 = 5,  /* transform + Wiener + tile aggregation */
    CPIPE_BM3D_T_MERGE2 = 6,  /* tile merge + normalisation + inverse colour */
    CPIPE_BM3D_T_COUNT = 7
};

typedef struct cpipe_bm3d_plan_s cpipe_bm3d_plan;

typedef struct {
    uint32_t struct_size; /* sizeof(cpipe_bm3d_plan_desc) */
    int32_t width;        /* pixels */
    int32_t height;
    int32_t channels;     /* 1 (grey plane) or 3 (R, G, B planes) */
    int32_t precision;    /* CPIPE_BM3D_PRECISION_* */
    int32_t device;       /* HIP device ordinal */
} cpipe_bm3d_plan_desc;

/* Last error message of the calling thread (never NULL). */
const char* cpipe_bm3d_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int cpipe_bm3d_device_count(void);

/* Plans own every device scratch buffer a run needs (allocated here, once). */
int cpipe_bm3d_plan_create(const cpipe_bm3d_plan_desc* desc, cpipe_bm3d_plan** out_plan);
void cpipe_bm3d_plan_destroy(cpipe_bm3d_plan* plan);

/* Both BM3D steps on device FP32 planes. in/out: `channels` plane pointers,
 * row pitch in elements (>= width). Inputs are read-only; out may not alias in.
 * sigma is the noise standard deviation of the input planes (>= 0). */
int cpipe_bm3d_plan_run(cpipe_bm3d_plan* plan, const float* const* in_planes, size_t in_pitch,
                        float* const* out_planes, size_t out_pitch, float sigma, void* stream);

/* BM3D on interleaved RGBA16F device images (strides in pixels): RGB denoised,
 * alpha passed through. Requires a plan with channels == 3. */
int cpipe_bm3d_plan_run_rgba16f(cpipe_bm3d_plan* plan, const uint16_t* in, size_t in_stride_px,
                                uint16_t* out, size_t out_stride_px, float sigma, void* stream);

/* Copies of the last run's intermediates (device destinations):
 * basic estimate planes in the input colour space, and the step-1/step-2
 * match lists (per reference position: 16 packed (y<<16|x) and a raw count). */
int cpipe_bm3d_plan_basic(cpipe_bm3d_plan* plan, float* const* out_planes, size_t out_pitch,
                          void* stream);
int cpipe_bm3d_plan_matches(cpipe_bm3d_plan* plan, int step, uint32_t* pos, uint8_t* count,
                            void* stream);
/* Synchronises the device and reports whether any run of this plan hit an
 * internal consistency check (CPIPE_INTERNAL_ERROR; results invalid). */
int cpipe_bm3d_plan_check(cpipe_bm3d_plan* plan);

/* Reference grid of the plan: number of reference positions along x and y. */
int cpipe_bm3d_plan_ref_grid(const cpipe_bm3d_plan* plan, int* nx, int* ny);

/* Per-stage HIP-event timing on the run stream (CPIPE_BM3D_T_*). */
int cpipe_bm3d_plan_set_timing(cpipe_bm3d_plan* plan, int enable);
int cpipe_bm3d_plan_timing(cpipe_bm3d_plan* plan, float* ms, int n, int* runs);

/* Reference-compatible filter (stage 1: 2x2 box-mean blend, stage 2:
 * identity) on interleaved RGBA16F device images, strides in pixels. */
int cpipe_bm3d_compat_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                              size_t out_stride_px, int width, int height, float sigma, int stage,
                              void* stream);

/* Host-memory conveniences used by the node: copy in, run, copy out, sync. */
int cpipe_bm3d_host_planes(const float* const* in_planes, size_t in_pitch, float* const* out_planes,
                           size_t out_pitch, int width, int height, int channels, float sigma,
                           int precision, int device);
int cpipe_bm3d_host_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                            size_t out_stride_px, int width, int height, float sigma, int mode,
                            int precision, int device);

#ifdef __cplusplus
}
#endif
#endif
