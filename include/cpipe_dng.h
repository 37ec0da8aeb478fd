/*
 * DNG NoiseProfile ingest for the BM3D node's noise-profile sigma map.
 *
 * The reference's ingest never fills CalibrationBlock::noise_profile
 * (include/cpipe/core/CalibrationBlock.hpp:54): src/cpipe/ingest/dng/DngReader.cpp
 * copies the calibration that src/cpipe/ingest/dng_opcode/OpcodeListParser.cpp
 * parsed from the TIFF tags (its tag list, OpcodeListParser.cpp:27-60, has no
 * NoiseProfile), so the node's metadata sigma (src/cpipe/nodes/denoise/bm3d.cpp:41-54)
 * always falls back to 0.03. This parser reads the DNG NoiseProfile tag (51041,
 * DOUBLE, [S0, O0, S1, O1, ...]: noise variance S * x + O per colour plane, or one
 * pair for all planes) from the bytes of a DNG file, so a host's DngReader can
 * fill noise_profile with pairs (beta, lambda) = (O, S), the order bm3d.cpp:50
 * reads them in.
 */
#ifndef CPIPE_DNG_H
#define CPIPE_DNG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads the NoiseProfile of a DNG (classic TIFF, either byte order) held in
 * memory. Walks IFD0, its chain and every SubIFD (tag 330); a NoiseProfile in
 * the main image IFD (NewSubFileType 0, the raw data) wins over one in a
 * preview IFD. On CPIPE_OK (0) *n_pairs is the number of pairs in the file (0
 * when there is no NoiseProfile) and the first min(n, max_pairs) are written
 * as beta[i] = O_i, lambda[i] = S_i (rounded to float). Returns
 * CPIPE_BAD_INDEX (5) for bad arguments and CPIPE_FAILED (1) for bytes that are
 * not a TIFF or whose NoiseProfile is malformed (wrong type, odd or zero count,
 * non-finite values, data outside the file). Never reads outside
 * [bytes, bytes + size). */
int cpipe_dng_noise_profile(const void* bytes, size_t size, float* beta, float* lambda, int max_pairs,
                            int* n_pairs);

#ifdef __cplusplus
}
#endif
#endif
