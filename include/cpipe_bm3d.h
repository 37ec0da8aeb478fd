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
    CPIPE_BM3D_PRECISION_BF16 = 1, /* BF16 MFMA group transforms, FP32 shrinkage and aggregation
                                      (+ optional MFMA block-matching cross term, FP32 re-rank) */
    CPIPE_BM3D_PRECISION_FP16 = 2  /* the same with FP16 operands */
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
    CPIPE_BM3D_T_PILOT = 7,   /* learned pilot + spot checks (replaces step 1) */
    CPIPE_BM3D_T_COUNT = 8
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
/* Waits for the plan's last run (an event on that run's stream, not the whole
 * device) and reports whether that run hit an internal consistency check
 * (CPIPE_INTERNAL_ERROR; its results are invalid). Every run clears the flag
 * first, so one failed run does not condemn the later ones. OK before any run. */
int cpipe_bm3d_plan_check(cpipe_bm3d_plan* plan);
/* Diagnostics for tests: marks the plan's last run as failed on `stream`, as an
 * internal consistency check would, so the next cpipe_bm3d_plan_check reports
 * CPIPE_INTERNAL_ERROR until the plan runs again. No counterpart in the
 * reference. */
int cpipe_bm3d_plan_debug_flag(cpipe_bm3d_plan* plan, void* stream);

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

/* ---------------------------------------------------------------------------
 * Learned pilot: a 6-layer 3x3 convolutional network (32 channels, FP16
 * weights and activations, FP32 accumulation, residual output) mapping the
 * noisy opponent planes and the noise level to the step-1 basic estimate.
 * No counterpart in the reference (build-only component of the node).
 * Planes are FP32 opponent-space Y, U, V; image i of a stack starts at
 * plane[c] + i * img_stride.
 * ------------------------------------------------------------------------- */
typedef struct cpipe_bm3d_pilot_s cpipe_bm3d_pilot;

const char* cpipe_bm3d_pilot_last_error(void);
size_t cpipe_bm3d_pilot_param_count(void);
int cpipe_bm3d_pilot_create(int device, cpipe_bm3d_pilot** out);
void cpipe_bm3d_pilot_destroy(cpipe_bm3d_pilot* net);
/* He-normal initialisation from a seed (deterministic). */
int cpipe_bm3d_pilot_init(cpipe_bm3d_pilot* net, uint64_t seed);
/* Versioned weight blob (magic "CPPL", version 1); with_optimizer adds the
 * Adam moments and step count (a training checkpoint). */
size_t cpipe_bm3d_pilot_blob_size(int with_optimizer);
int cpipe_bm3d_pilot_save(const cpipe_bm3d_pilot* net, void* blob, size_t cap, int with_optimizer);
int cpipe_bm3d_pilot_load(cpipe_bm3d_pilot* net, const void* blob, size_t size);
/* Inference on device planes (row pitch `pitch`), on the network's own
 * activation workspace: one stream at a time per network (plans that use the
 * network bring their own workspace, see cpipe_bm3d_plan_set_pilot). */
int cpipe_bm3d_pilot_forward(cpipe_bm3d_pilot* net, const float* const* in, size_t pitch,
                             size_t img_stride, int width, int height, int n_img, float sigma,
                             float* const* out, void* stream);
/* Training primitives on a stack of n_img images laid out [img][c][plane]
 * (in[c] = base + c * width * height, image stride 3 * width * height), with
 * one sigma per image (host array). Loss = 0.5 * mean squared error over
 * pixels and channels. gradients: forward + backward only (FP32 dLoss/dparam
 * copied to grads_host, cpipe_bm3d_pilot_param_count() floats); train_step:
 * forward + backward + one Adam update. */
int cpipe_bm3d_pilot_gradients(cpipe_bm3d_pilot* net, const float* const* in,
                               const float* const* target, int width, int height, int n_img,
                               const float* sigma, float* grads_host, float* loss);
/* Training loss: mode 0 (default) = 1/2 MSE against the targets; mode 1 =
 * the same with image i weighted by (ref_sigma / sigma_i)^2, i.e. the error
 * measured against each crop's noise energy, the quantity the spot checks'
 * method-noise ratio tests. Applies to _gradients and _train_step. */
int cpipe_bm3d_pilot_set_loss(cpipe_bm3d_pilot* net, int mode, float ref_sigma);
int cpipe_bm3d_pilot_train_step(cpipe_bm3d_pilot* net, const float* const* in,
                                const float* const* target, int width, int height, int n_img,
                                const float* sigma, float lr, float* loss);

/* Step 1 only, for training the pilot: writes the opponent-space noisy planes
 * and the plain basic estimate (3 planes each, device, row pitch out_pitch). */
int cpipe_bm3d_plan_step1_opponent(cpipe_bm3d_plan* plan, const float* const* in, size_t in_pitch,
                                   float sigma, float* const* noisy_out, float* const* basic_out,
                                   size_t out_pitch, void* stream);

/* Step 2 only, on a given basic estimate: the Wiener step of a plain run with
 * the opponent-space basic estimate basic_in (3 device planes, row pitch
 * basic_pitch) in place of step 1's, writing the denoised RGB planes to out.
 * With the basic planes of cpipe_bm3d_plan_step1_opponent it reproduces
 * cpipe_bm3d_plan_run; it scores a candidate pilot network by the output it
 * leads to. */
int cpipe_bm3d_plan_step2_opponent(cpipe_bm3d_plan* plan, const float* const* in, size_t in_pitch,
                                   float sigma, const float* const* basic_in, size_t basic_pitch,
                                   float* const* out, size_t out_pitch, void* stream);

/* Learned pilot in a plan (3-channel plans): step 1 is replaced by the
 * network; blocks of check_block x check_block pixels, one in check_every
 * (block b = by * nbx + bx in raster order is checked when
 * (b + by) % check_every == check_every / 2), are spot-checked against the
 * plain step 1 computed exactly on that block. With the method-noise ratio
 * r = 10 log10(|z - pilot|^2 / |z - plain|^2) (dB), a block uses the plain
 * estimate when r > fallback_db (the pilot removes more than plain step 1:
 * over-smoothing) or r < -under_db (it leaves noise in: the Wiener step then
 * keeps it, so the bound is tighter); if more than full_fraction of the
 * checked blocks fall back, the plain step 1 runs on the whole frame (decided
 * on the device, no host sync). The checks compare the network's output on
 * each checked block from its own run over the block and a 6-pixel border
 * (the same values as the whole-frame run there), so the decision can come
 * while the whole-frame network still runs; once a frame falls back, the
 * network's remaining layers skip their work. net = NULL restores the plain
 * path. The plan
 * keeps a reference to net (which must outlive its use) and allocates its own
 * activation workspace for its largest frame, so plans sharing one network
 * may run on different streams at once. A struct_size that ends before
 * under_db selects under_db = fallback_db / 2. */
typedef struct {
    uint32_t struct_size;
    float fallback_db;    /* default 0.5 */
    int32_t check_block;  /* default 256 */
    int32_t check_every;  /* default 16; 0 disables spot checks */
    float full_fraction;  /* whole-frame fallback when more than this fraction of the
                             checked blocks fall back; default 0.5 (>= 1 disables it) */
    float under_db;       /* default 0.25 */
} cpipe_bm3d_pilot_opts;
int cpipe_bm3d_plan_set_pilot(cpipe_bm3d_plan* plan, cpipe_bm3d_pilot* net,
                              const cpipe_bm3d_pilot_opts* opts);
/* Outcome of the last run's spot checks (synchronises the device). */
int cpipe_bm3d_plan_pilot_report(cpipe_bm3d_plan* plan, int* n_checked, int* n_fallback,
                                 float* max_delta_db, int* full_fallback);
/* (full_fallback counts the frames whose gate ran the plain step 1 on the whole frame.) */

/* ---------------------------------------------------------------------------
 * Batches: frames of uneven sizes in one plan. Every stage is one launch over
 * a work list spanning all frames (block-matching tiles, aggregation tiles,
 * pixel bands), each frame with its own sigma; with use_graph the whole batch
 * is captured once as a hipGraph and each run is one graph launch. No
 * counterpart in the reference (its node processes one buffer per call,
 * bm3d.cpp:80-104); the per-frame semantics equal cpipe_bm3d_plan_run.
 * A batch is a plan: cpipe_bm3d_plan_set_pilot, _set_timing, _timing, _check
 * and _pilot_report apply to it.
 * ------------------------------------------------------------------------- */
typedef struct cpipe_bm3d_plan_s cpipe_bm3d_batch;

typedef struct {
    uint32_t struct_size;    /* sizeof(cpipe_bm3d_batch_desc) */
    int32_t n_frames;        /* 1 .. 4096 */
    const int32_t* widths;   /* n_frames entries */
    const int32_t* heights;
    int32_t channels;        /* 1 or 3, all frames */
    int32_t precision;       /* CPIPE_BM3D_PRECISION_* */
    int32_t device;
    int32_t use_graph;       /* 1: replay one captured hipGraph per run */
} cpipe_bm3d_batch_desc;

int cpipe_bm3d_batch_create(const cpipe_bm3d_batch_desc* desc, cpipe_bm3d_batch** out_batch);
void cpipe_bm3d_batch_destroy(cpipe_bm3d_batch* batch);
/* in/out: n_frames * channels device plane pointers, frame-major; pitches (elements)
 * and sigmas per frame. Asynchronous on `stream`. */
int cpipe_bm3d_batch_run(cpipe_bm3d_batch* batch, const float* const* in_planes,
                         const size_t* in_pitch, float* const* out_planes, const size_t* out_pitch,
                         const float* sigmas, void* stream);

/* Autotuner: times every candidate launch configuration on these frames
 * (`iters` runs each, after one warm-up), keeps the fastest for the batch and
 * caches it for later batches of the same shape (frame sizes, channels,
 * precision). Candidates: aggregation tile height and block-matching kernel
 * (mixed precision only: 8 or 16 refs, MFMA or exact block matching; exact
 * FP32 keeps 16 x 4-ref tiles, the CPU restatement's tiling, and exact
 * matching),
 * work-list order (frame raster, or one contiguous run per XCD) and pixel-band
 * height.
 * Synchronises the stream. */
#define CPIPE_BM3D_TUNE_MAX 32
typedef struct {
    int32_t n, best;
    int32_t tile_y[CPIPE_BM3D_TUNE_MAX];
    int32_t order[CPIPE_BM3D_TUNE_MAX];
    int32_t band_rows[CPIPE_BM3D_TUNE_MAX];
    int32_t bm_mfma[CPIPE_BM3D_TUNE_MAX];
    float ms[CPIPE_BM3D_TUNE_MAX];
} cpipe_bm3d_tune_report;
int cpipe_bm3d_batch_tune(cpipe_bm3d_batch* batch, const float* const* in_planes,
                          const size_t* in_pitch, float* const* out_planes, const size_t* out_pitch,
                          const float* sigmas, int iters, void* stream,
                          cpipe_bm3d_tune_report* report);
/* Current launch configuration of a plan or batch. */
int cpipe_bm3d_batch_config(const cpipe_bm3d_batch* batch, int* tile_y, int* order, int* band_rows,
                            int* bm_mfma);

/* Mixed-precision plans: block matching with the FP16/BF16 MFMA cross term and
 * an exact FP32 re-rank of a 48-candidate shortlist per ref (mfma = 1, the
 * default), or the exact FP32 kernel (mfma = 0). Both produce the exact kernel's
 * match lists whenever the shortlist holds the exact best 15 (always, on the
 * test images); the autotuner times both. CPIPE_BM3D_UNSUPPORTED for FP32 plans. */
int cpipe_bm3d_plan_set_block_matching(cpipe_bm3d_plan* plan, int mfma);

/* Noise-profile sigma map (no counterpart in the reference, whose node reduces
 * the DNG NoiseProfile pairs of CalibrationBlock::noise_profile to one sigma,
 * src/cpipe/nodes/denoise/bm3d.cpp:21-55). With n = channels pairs, later runs
 * of a single-frame plan give each reference position and (opponent) channel
 * its own sigma: variance v_c = max(0, beta[c] + lambda[c] * x_c) at the mean
 * x_c of the reference's 8x8 input patch, opponent channel k sigma^2 =
 * sum_c A_kc^2 v_c (A the opponent rows (1/3, 1/3, 1/3), (1/2, 0, -1/2),
 * (1/4, -1/2, 1/4)); each group shrinks with its reference's sigma and its
 * aggregation weight is divided by sigma^2. The runs' sigma argument is then
 * unused. n = 0 returns to the scalar sigma. On a batch the pairs apply to
 * every frame. CPIPE_BM3D_UNSUPPORTED together with the learned pilot. */
int cpipe_bm3d_plan_set_noise_profile(cpipe_bm3d_plan* plan, const float* beta, const float* lambda,
                                      int n);
/* The same per frame of a batch: beta/lambda hold n_frames x channels values
 * (frame-major); pairs_per_frame = channels, or 0 to return to the sigmas. */
int cpipe_bm3d_batch_set_noise_profiles(cpipe_bm3d_batch* batch, const float* beta,
                                        const float* lambda, int pairs_per_frame);

/* Host-memory conveniences used by the node (replace the Halide dispatch of
 * bm3d_dispatch.cpp:19-41 behind DenoiseBm3d::process, bm3d.cpp:96-100):
 * copy in, run, copy out, synchronise. The reference host calls a node's
 * create/process/destroy on every run (Pipeline.cpp:716-772), so the plans,
 * device buffers and the learned pilot are kept in a process-global,
 * mutex-protected cache (the last CPIPE_BM3D_HOST_SLOTS shapes, least recently
 * used evicted) and reused across calls; calls are serialised by the cache's
 * lock. flags: CPIPE_BM3D_HOST_PILOT runs the learned pilot (weights embedded
 * in the library) in place of the plain step 1 (3-channel images only). */
#define CPIPE_BM3D_HOST_PILOT 1
/* flags: CPIPE_BM3D_HOST_PROFILE runs with the noise-profile sigma map of the
 * pairs last given to cpipe_bm3d_host_noise_profile on the calling thread (one
 * pair for every channel, or one per channel); mode bm3d, not with the pilot. */
#define CPIPE_BM3D_HOST_PROFILE 2
#define CPIPE_BM3D_HOST_SLOTS 4
int cpipe_bm3d_host_noise_profile(const float* beta, const float* lambda, int n);
int cpipe_bm3d_host_planes(const float* const* in_planes, size_t in_pitch, float* const* out_planes,
                           size_t out_pitch, int width, int height, int channels, float sigma,
                           int precision, int device, int flags);
int cpipe_bm3d_host_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                            size_t out_stride_px, int width, int height, float sigma, int mode,
                            int precision, int device, int flags);
/* The same on device-resident images (the node's path when the host keeps its
 * buffers in HIP device memory; replaces the same dispatch with the copies and
 * the lock_cpu round trip, HalideBufferAdapter.cpp:41, taken out). Pointers are
 * device addresses on `device`, pitches/strides as above; work is ordered on
 * `stream` (hipStream_t as void*, NULL = null stream) and the call returns once
 * the output is written and the plan's consistency flag is read. Same plan
 * cache and lock as the host conveniences; no staging buffers are allocated. */
int cpipe_bm3d_device_planes(const float* const* in_planes, size_t in_pitch,
                             float* const* out_planes, size_t out_pitch, int width, int height,
                             int channels, float sigma, int precision, int device, int flags,
                             void* stream);
int cpipe_bm3d_device_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                              size_t out_stride_px, int width, int height, float sigma, int mode,
                              int precision, int device, int flags, void* stream);

/* BM3D without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The same run as
 * cpipe_bm3d_device_planes / cpipe_bm3d_plan_run with precision FP32 and the
 * plans' defaults (exact FP32, both steps, no learned pilot, no noise-profile
 * map), with bit-identical output, on device planes of the current HIP device
 * (pitches in floats), enqueued on `stream` (hipStream_t as void*, NULL = null
 * stream) and not waited for.
 *
 * Every buffer the run needs lives in one static device array of
 * CPIPE_BM3D_NOALLOC_SCRATCH_MIB MiB that each GPU the library runs on holds
 * until the process ends; the call itself calls no hipMalloc, hipFree or other
 * allocating routine and copies nothing from the host. Frames up to
 * CPIPE_BM3D_NOALLOC_MAX_PIXELS pixels with both sides up to
 * CPIPE_BM3D_NOALLOC_MAX_SIDE are accepted, 1 or 3 channels; larger frames are
 * refused with CPIPE_BM3D_BAD_INDEX. Calls on one GPU share its array: calls
 * that overlap in time there (other streams, or another thread) race, so order
 * them on one stream or synchronise between them. The first call on a device
 * looks up the array's address and uploads the constant tables (keep it out of
 * a graph capture); later calls enqueue kernels only. The aggregation's
 * consistency flag is cleared and set on the device but not read back. */
#define CPIPE_BM3D_NOALLOC_SCRATCH_MIB 320
#define CPIPE_BM3D_NOALLOC_MAX_PIXELS 1048576
#define CPIPE_BM3D_NOALLOC_MAX_SIDE 8192
int cpipe_bm3d_noalloc_planes(const float* const* in_planes, size_t in_pitch,
                              float* const* out_planes, size_t out_pitch, int width, int height,
                              int channels, float sigma, void* stream);
/* Frees every cached plan, buffer and pilot network of the host conveniences. */
void cpipe_bm3d_host_release(void);
/* The learned-pilot weight blob embedded in the library (cpipe_amd/weights/pilot_v1.bin). */
const void* cpipe_bm3d_pilot_embedded(size_t* size);

#ifdef __cplusplus
}
#endif
#endif
