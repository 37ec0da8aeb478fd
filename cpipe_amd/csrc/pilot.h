// Internal C++ interface between the BM3D plan and the learned pilot.
#pragma once
#include <hip/hip_runtime.h>

#include "cpipe_bm3d.h"

namespace cpipe_internal {
// Runs the network on n_img stacked images of w x h opponent planes
// (in[c] + img * img_stride + y * pitch + x) and writes the pilot planes.
// `levels` holds one noise level (10 * sigma) per image on the device, or is
// null to use `level` for all. The activations go to `ws`, the caller's device
// workspace of at least pilot_workspace_bytes(w * h * n_img) bytes, so that
// callers sharing one network on different streams, or capturing its launches
// into a graph, never share or lose activation memory; ws = null uses the
// network's own workspace (grown on demand; one stream at a time).
size_t pilot_workspace_bytes(size_t px);
// `gate` (device int, may be null): every layer skips its work once it reads a
// nonzero value there (a frame whose spot checks fell back discards the output).
int pilot_forward(cpipe_bm3d_pilot* net, const float* const* in, size_t pitch, size_t img_stride,
                  int w, int h, int n_img, float level, const float* levels,
                  float* const* out, void* ws, hipStream_t st, const int* gate = nullptr);

}  // namespace cpipe_internal
