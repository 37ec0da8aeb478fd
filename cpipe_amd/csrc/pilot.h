// Internal C++ interface between the BM3D plan and the learned pilot.
#pragma once
#include <hip/hip_runtime.h>

#include "cpipe_bm3d.h"

namespace cpipe_internal {
// Runs the network on n_img stacked images of w x h opponent planes
// (in[c] + img * img_stride + y * pitch + x) and writes the pilot planes.
// `levels` holds one noise level (10 * sigma) per image on the device, or is
// null to use `level` for all.
int pilot_forward(cpipe_bm3d_pilot* net, const float* const* in, size_t pitch, size_t img_stride,
                  int w, int h, int n_img, float level, const float* levels,
                  float* const* out, hipStream_t st);

}  // namespace cpipe_internal
