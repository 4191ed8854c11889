// pds_binned.hip -- tile-binned forward (placeholder: photon-order flat path).
#include "pds_binned.hpp"

namespace pds {

static thread_local const char* g_err = "";

const char* binned_last_error() { return g_err; }

size_t binned_workspace_bytes(int64_t, int32_t, int32_t, int32_t, int32_t, int32_t, const pds_tuning*) { return 0; }

int binned_forward(const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning*, float* out, void*,
                   size_t, hipStream_t s) {
    if (mode != PDS_MODE_AUTO) {
        g_err = "binned modes are not built yet";
        return PDS_ERR_ARGUMENT;
    }
    if (launch_fill_zero(out, (int64_t)I.C * I.H * I.W, s) != hipSuccess) return PDS_ERR_HIP;
    if (launch_forward_flat(P, I, out, s) != hipSuccess) return PDS_ERR_HIP;
    return PDS_OK;
}

}  // namespace pds
