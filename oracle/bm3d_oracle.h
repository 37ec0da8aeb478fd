/*
 * CPU oracle for the BM3D denoise node -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the reference result (or the timed CPU baseline).
 * The product path (cpipe_amd/) never links or calls it.
 *
 * Restatements (the BM3D node's three, the wavelet node's one):
 *   1. oracle_compat_rgba16f  -- the reference node's actual arithmetic:
 *      src/cpipe/nodes/denoise/bm3d_step1_generator.cpp:9-16,42-56 (step 1,
 *      2x2 box-mean blend) and bm3d_step2_generator.cpp:33-42 (identity).
 *   2. oracle_bm3d_f64        -- textbook two-step BM3D in FP64, single
 *      threaded, direct loops (quality reference for PSNR / max-abs).
 *      The reference ships no BM3D arithmetic; the algorithm follows its
 *      design prose docs/research/07-classic-isp-algorithms.md:277-283,541-560
 *      and the public papers (Dabov et al. 2007; Lebrun, IPOL 2012).
 *   3. oracle_bm3d_f32        -- the same algorithm in FP32 with exactly the
 *      operation order of the GPU FP32 path (fma chains, tie rules,
 *      aggregation tiles and merge order), multithreaded. It is both the
 *      parity partner of the GPU path and the CPU baseline. Variants
 *      with a per-reference sigma from a noise profile (oracle_sigma_map_f32,
 *      oracle_bm3d_f32_profile / _f64_profile).
 * and, for the downstream node com.cpipe.denoise.wavelet_bayes_shrink,
 *   4. oracle_wavelet_rgba16f -- its Halide generator's arithmetic
 *      (wavelet_oracle.c);
 * and for com.cpipe.precision_convert, oracle_convert_* (convert_oracle.c).
 *
 * Parity pins (see tests/test_oracle_*.py): compat known answers from the
 * reference unit tests; DCT/Haar orthonormality; Kaiser from the I0 series;
 * sigma->0 identity; constant image; brute-force block matching on tiny
 * inputs; FP32 mirror vs FP64 textbook agreement.
 */
#ifndef CPIPE_BM3D_ORACLE_H
#define CPIPE_BM3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Algorithm constants (node defaults). */
#define ORC_K 8        /* patch edge */
#define ORC_R 19       /* search radius: 39x39 window */
#define ORC_STRIDE 3   /* reference-patch stride, last row/col forced */
#define ORC_NMAX 16    /* maximum group size */
#define ORC_TILE_X 16  /* refs per aggregation tile along x (GPU tiling, mirrored) */
#define ORC_TILE_Y 4   /* refs per aggregation tile along y */

/* IEEE binary16 <-> binary32, round to nearest even. */
uint16_t oracle_f32_to_f16(float v);
float oracle_f16_to_f32(uint16_t h);

/* Reference-compatible node arithmetic on interleaved RGBA16F.
 * stage 1: 2x2 box-mean blend; stage 2: identity. Strides in pixels. */
void oracle_compat_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                           size_t out_stride_px, int width, int height, float sigma, int stage);

/* Number of reference positions along an axis of n pixels, and the i-th. */
int oracle_ref_count(int n);
int oracle_ref_pos(int i, int n);

/* 8x8 orthonormal DCT-II matrix (row u = frequency) and 8-tap Kaiser (beta=2). */
void oracle_dct_matrix_f64(double c[64]);
/* One 8-point orthonormal DCT-II (inverse = 1: its inverse) in FP32 with the
 * butterfly order the FP32 BM3D restatement and the GPU exact kernels use. */
void oracle_dct8_f32(const float* in, float* out, int inverse);
void oracle_kaiser_f64(double k[8]);

/* Block matching in FP32 on one plane. Writes for every reference position
 * (row-major over the ref grid) up to 16 packed positions (y<<16 | x) and the
 * raw match count (1..16, before power-of-two rounding). tau_ssd is the SSD
 * threshold (sum over the 64 pixels, not normalised). */
void oracle_bm_f32(const float* plane, int width, int height, float tau_ssd, uint32_t* pos,
                   uint8_t* count, int nthreads);
void oracle_bm_f64(const double* plane, int width, int height, double tau_ssd, uint32_t* pos,
                   uint8_t* count);

/* Full BM3D, FP32 mirror of the GPU path. channels = 1 or 3 (RGB planes; the
 * opponent transform is applied inside). basic (nullable) receives the step-1
 * estimate in the input colour space. matches1/counts1/matches2/counts2 are
 * nullable debug outputs (sized nref*16 and nref). Returns 0 on success. */
int oracle_bm3d_f32(const float* const* in, float* const* out, float* const* basic, int width,
                    int height, int channels, float sigma, int nthreads, uint32_t* matches1,
                    uint8_t* counts1, uint32_t* matches2, uint8_t* counts2);

/* Full BM3D, FP64 textbook. Same I/O conventions (double planes). */
int oracle_bm3d_f64(const double* const* in, double* const* out, double* const* basic,
                    int width, int height, int channels, double sigma);

/* Noise-profile sigma map: per reference position and (opponent) channel, the
 * sigma of variance beta[c] + lambda[c] * x at the reference patch's mean x
 * (smap: channels x nref floats, channel-major; `in` planes with row pitch
 * `pitch`). The BM3D variants below use it in place of the scalar sigma in the
 * group shrinkage and multiply each group's aggregation weight by 1/sigma^2
 * (the textbook weights 1/(sigma^2 N) and 1/(sigma^2 sum W^2)). */
int oracle_sigma_map_f32(const float* const* in, size_t pitch, int width, int height, int channels,
                         const float* beta, const float* lambda, float* smap);
int oracle_bm3d_f32_profile(const float* const* in, float* const* out, float* const* basic,
                            int width, int height, int channels, const float* beta,
                            const float* lambda, int nthreads);
int oracle_bm3d_f64_profile(const double* const* in, double* const* out, double* const* basic,
                            int width, int height, int channels, const double* beta,
                            const double* lambda);

/* Wavelet BayesShrink chroma denoise of interleaved RGBA16F (wavelet_oracle.c):
 * the reference Halide generator denoise_wavelet_bayes_shrink
 * (src/cpipe/nodes/denoise/wavelet_bayes_shrink_generator.cpp:57-86). Strides in pixels. */
void oracle_wavelet_rgba16f(const uint16_t* in, size_t in_stride_px, uint16_t* out,
                            size_t out_stride_px, int width, int height, float chroma_strength);

/* Precision conversions of the precision-convert node (convert_oracle.c): the
 * reference Halide generators src/cpipe/nodes/precision/precision_convert_generator.cpp.
 * Strides in bytes; bits = 8 (R8G8B8A8_UNORM) or 16 (R16G16B16A16_UNORM). */
void oracle_convert_r16u_to_f32(const uint8_t* in, size_t in_stride, uint8_t* out, size_t out_stride,
                                int width, int height);
void oracle_convert_f32_to_rgba16f(const uint8_t* in, size_t in_stride, uint8_t* out,
                                   size_t out_stride, int width, int height);
void oracle_convert_rgba16f_to_unorm(const uint8_t* in, size_t in_stride, uint8_t* out,
                                     size_t out_stride, int width, int height, int bits);

#ifdef __cplusplus
}
#endif
#endif
