// pds_binned.hpp -- the tile-binned forward (pds_binned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pds.h"
#include "pds_internal.hpp"

namespace pds {

// Workspace of the binned path.  keep: the buffer also holds what a later
// backward reuses (record slots per photon, photon classes, partial sums).
size_t binned_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t cap, int32_t mode,
                              const pds_tuning* t, bool ell_trivial, bool keep);
// Forward (any non-plain mode).  keep: leave the binning in ws for binned_backward(reuse).
int binned_forward(const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning* t, float* out,
                   void* ws, size_t ws_bytes, bool keep, hipStream_t s);
// Backward through the bins; reuse: ws holds the binning of a keep-forward of the same inputs.
int binned_backward(const float* grad, const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning* t,
                    const GradArgs& G, void* ws, size_t ws_bytes, bool reuse, hipStream_t s);
// Statistics of a kept binning (synchronises s).
int binned_stats(const void* ws, size_t ws_bytes, int64_t N, const ImageArgs& I, int32_t mode, const pds_tuning* t,
                 pds_bins_stats* out, hipStream_t s);
bool binned_possible(const ImageArgs& I, int64_t N, int32_t mode, const pds_tuning* t);
const char* binned_last_error();

}  // namespace pds
