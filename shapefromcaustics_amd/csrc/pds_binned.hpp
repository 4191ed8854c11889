// pds_binned.hpp -- the tile-binned forward (pds_binned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pds.h"
#include "pds_internal.hpp"

namespace pds {

size_t binned_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t cap, int32_t mode,
                              const pds_tuning* t, bool ell_trivial);
int binned_forward(const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning* t, float* out,
                   void* ws, size_t ws_bytes, hipStream_t s);
const char* binned_last_error();

}  // namespace pds
