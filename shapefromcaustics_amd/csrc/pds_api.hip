// pds_api.hip -- the C ABI of libpds_hip.so (declared in include/pds.h).
//
// Replaces the reference's pybind entry points pds_forward / pds_backward
// (PyOptix/PhotonDifferentialSplattig.cpp:12-32) and their CUDA launchers
// (PyOptix/kernel/photon_differentials.cu:115-134, :223-247) with plain
// pointers, caller-owned buffers and an explicit stream.  Validation of
// tensor-level properties (device, contiguity, the reference's shape
// messages) happens in the Python binding; this layer validates sizes,
// pointers, the mode and the workspace, and reports every failure through a
// status code plus pds_last_error().
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "pds.h"
#include "pds_binned.hpp"
#include "pds_internal.hpp"

namespace {

thread_local char g_last_error[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(PDS_ERR_HIP, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
}

// Host twin of pds::window_ellipse_ok: the reference's window ellipse test
// (photon_differentials.cu:93) is monotone in |x_off| and |y_off|, so if it
// holds at (cap, cap) it holds for every offset of every window.
bool ellipse_trivial(int32_t cap, int32_t W, int32_t H) {
    if (cap < 0) return true;
    const double w = (double)(float)W, h = (double)(float)H;
    const double a = (double)(float)((int64_t)cap * cap) / (0.25 * w * w);
    const double b = (double)(float)((int64_t)cap * cap) / (0.25 * h * h);
    return a + b <= 1.0;
}

int check_common(int64_t N, int32_t C, int32_t H, int32_t W, int32_t mode, const void* E, const void* xp,
                 const void* M, const void* cp, const void* radius) {
    if (N < 0) return fail(PDS_ERR_ARGUMENT, "photon count N = %lld is negative", (long long)N);
    if (C < 0 || H < 0 || W < 0)
        return fail(PDS_ERR_ARGUMENT, "output size (%d, %d, %d) has a negative extent", C, H, W);
    if ((int64_t)C * H * W >= ((int64_t)1 << 31))
        return fail(PDS_ERR_CAPACITY, "output of %lld elements exceeds the 2^31 - 1 element limit",
                    (long long)C * H * W);
    if (mode < PDS_MODE_AUTO || mode > PDS_MODE_DETERMINISTIC) return fail(PDS_ERR_ARGUMENT, "unknown mode %d", mode);
    if (N > 0 && (!E || !xp || !M || !cp || !radius))
        return fail(PDS_ERR_ARGUMENT, "null photon input pointer with N = %lld", (long long)N);
    return PDS_OK;
}

pds::PhotonArgs photon_args(const float* E, const float* xp, const float* M, const int64_t* cp, const float* radius,
                            int64_t N, const int64_t* live = nullptr) {
    pds::PhotonArgs P;
    P.live = live;
    P.E = E;
    P.xp = xp;
    P.M = M;
    P.cp = cp;
    P.radius = radius;
    P.N = N;
    return P;
}

size_t bins_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t cap, int32_t mode, const pds_tuning* t) {
    if (N < 0 || C <= 0 || H <= 0 || W <= 0 || mode == PDS_MODE_PLAIN) return 0;
    return pds::binned_workspace_bytes(N, C, H, W, cap, mode, t, ellipse_trivial(cap, W, H), /*keep=*/true);
}

pds::ImageArgs image_args(int32_t C, int32_t H, int32_t W, int32_t cap) {
    pds::ImageArgs I;
    I.C = C;
    I.H = H;
    I.W = W;
    I.cap = cap;
    I.fw = (float)W;
    I.fh = (float)H;
    I.inv_hw = (float)(2.0 / (double)W);
    I.inv_hh = (float)(2.0 / (double)H);
    I.ell_trivial = ellipse_trivial(cap, W, H) ? 1 : 0;
    return I;
}

}  // namespace

namespace {

// live: null, or a device photon count (kept bins only)
int forward_impl(const float* Ep, const float* xp, const float* Mp, const int64_t* cp, const float* radius,
                 int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius, int32_t mode,
                 const pds_tuning* t, float* out, void* workspace, size_t workspace_bytes, bool keep, void* stream,
                 const int64_t* live = nullptr) {
    g_last_error[0] = '\0';
    int rc = check_common(N, C, H, W, mode, Ep, xp, Mp, cp, radius);
    if (rc) return rc;
    const int64_t numel = (int64_t)C * H * W;
    if (numel == 0) return PDS_OK;
    if (!out) return fail(PDS_ERR_ARGUMENT, "null output pointer");
    if (keep && mode == PDS_MODE_PLAIN) return fail(PDS_ERR_ARGUMENT, "the plain mode keeps no bins");
    const size_t need = keep ? bins_bytes(N, C, H, W, max_pixel_radius, mode, t)
                             : pds_forward_workspace_bytes(N, C, H, W, max_pixel_radius, mode, t);
    if (keep && need == 0) return fail(PDS_ERR_ARGUMENT, "this problem cannot be binned: use pds_forward_f32");
    if (need > 0 && (workspace == nullptr || workspace_bytes < need))
        return fail(PDS_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required",
                    workspace ? workspace_bytes : (size_t)0, need);
    hipStream_t s = (hipStream_t)stream;
    const pds::PhotonArgs P = photon_args(Ep, xp, Mp, cp, radius, N, live);
    const pds::ImageArgs I = image_args(C, H, W, max_pixel_radius);
    hipError_t e;
    if (mode == PDS_MODE_PLAIN) {
        if ((e = pds::launch_fill_zero(out, numel, s)) != hipSuccess) return hip_fail(e, "zero fill");
        if ((e = pds::launch_forward_plain(P, I, out, s)) != hipSuccess) return hip_fail(e, "plain forward");
        return PDS_OK;
    }
    rc = pds::binned_forward(P, I, mode, t, out, workspace, workspace_bytes, keep, s);
    if (rc == PDS_ERR_HIP) return hip_fail(hipGetLastError(), "binned forward");
    if (rc) return fail(rc, "binned forward: %s", pds::binned_last_error());
    return PDS_OK;
}

int backward_impl(const float* grad, const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                  const float* radius, int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                  int32_t mode, const pds_tuning* t, float* grad_Ep, float* grad_xp, float* grad_Mp, void* workspace,
                  size_t workspace_bytes, bool reuse, void* stream, const int64_t* live = nullptr) {
    g_last_error[0] = '\0';
    int rc = check_common(N, C, H, W, mode, Ep, xp, Mp, cp, radius);
    if (rc) return rc;
    if (N == 0) return PDS_OK;
    if (!grad_Ep || !grad_xp || !grad_Mp) return fail(PDS_ERR_ARGUMENT, "null gradient output pointer");
    const int64_t numel = (int64_t)C * H * W;
    if (numel > 0 && !grad) return fail(PDS_ERR_ARGUMENT, "null upstream gradient pointer");
    hipStream_t s = (hipStream_t)stream;
    const pds::PhotonArgs P = photon_args(Ep, xp, Mp, cp, radius, N, live);
    const pds::ImageArgs I = image_args(C, H, W, max_pixel_radius);
    const pds::GradArgs G = {grad_Ep, grad_xp, grad_Mp};
    hipError_t e;
    if (numel == 0) {
        // every photon is skipped: all gradients are zero
        if ((e = pds::launch_fill_zero(grad_Ep, N, s)) != hipSuccess) return hip_fail(e, "zero fill");
        if ((e = pds::launch_fill_zero(grad_xp, 2 * N, s)) != hipSuccess) return hip_fail(e, "zero fill");
        if ((e = pds::launch_fill_zero(grad_Mp, 6 * N, s)) != hipSuccess) return hip_fail(e, "zero fill");
        return PDS_OK;
    }
    const size_t need = bins_bytes(N, C, H, W, max_pixel_radius, mode, t);
    if (reuse && need == 0) return fail(PDS_ERR_ARGUMENT, "this problem has no bins: use pds_backward_f32");
    if (mode != PDS_MODE_PLAIN && need > 0) {
        if (workspace == nullptr || workspace_bytes < need)
            return fail(PDS_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required",
                        workspace ? workspace_bytes : (size_t)0, need);
        rc = pds::binned_backward(grad, P, I, mode, t, G, workspace, workspace_bytes, reuse, s);
        if (rc == PDS_ERR_HIP) return hip_fail(hipGetLastError(), "binned backward");
        if (rc) return fail(rc, "binned backward: %s", pds::binned_last_error());
        return PDS_OK;
    }
    if (mode == PDS_MODE_PLAIN || !I.ell_trivial) {
        if ((e = pds::launch_backward_plain(grad, P, I, G, s)) != hipSuccess) return hip_fail(e, "plain backward");
        return PDS_OK;
    }
    if ((e = pds::launch_backward_flat(grad, P, I, G, s)) != hipSuccess) return hip_fail(e, "backward");
    return PDS_OK;
}

// ---------------------------------------------------------------------------
// Scratch of pds_forward_noalloc_f32: one fixed array per GPU, part of this
// library's device code object, so that the entry point needs no allocation.
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address comes from hipGetSymbolAddress, once per device (each GPU
// holds its own copy at its own address), and only that address is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PDS_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address (resolved on its first use there), or
// null with the error recorded.
void* device_scratch() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        hip_fail(e, "hipGetDevice");
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        fail(PDS_ERR_ARGUMENT, "device %d is beyond the %d devices the static scratch supports", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

const char* pds_last_error(void) { return g_last_error; }

const char* pds_version(void) { return "pds_hip 0.1 gfx950"; }

size_t pds_forward_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                                   int32_t mode, const pds_tuning* t) {
    if (N < 0 || C <= 0 || H <= 0 || W <= 0 || mode == PDS_MODE_PLAIN) return 0;
    return pds::binned_workspace_bytes(N, C, H, W, max_pixel_radius, mode, t,
                                       ellipse_trivial(max_pixel_radius, W, H), /*keep=*/false);
}

size_t pds_bins_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius, int32_t mode,
                      const pds_tuning* t) {
    return bins_bytes(N, C, H, W, max_pixel_radius, mode, t);
}

size_t pds_backward_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                                    int32_t mode, const pds_tuning* t) {
    return pds_bins_bytes(N, C, H, W, max_pixel_radius, mode, t);
}

int pds_forward_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp, const float* radius,
                    int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius, int32_t mode,
                    const pds_tuning* t, float* out, void* workspace, size_t workspace_bytes, void* stream) {
    return forward_impl(Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, out, workspace,
                        workspace_bytes, false, stream);
}

int pds_forward_noalloc_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                            const float* radius, int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                            int32_t mode, const pds_tuning* t, float* out, void* stream) {
    g_last_error[0] = '\0';
    int rc = check_common(N, C, H, W, mode, Ep, xp, Mp, cp, radius);
    if (rc) return rc;
    const size_t need = pds_forward_workspace_bytes(N, C, H, W, max_pixel_radius, mode, t);
    if (need > PDS_STATIC_SCRATCH_BYTES)
        return fail(PDS_ERR_CAPACITY,
                    "pds_forward_noalloc_f32: this problem needs %zu bytes of scratch, more than the %zu bytes of the "
                    "static scratch (use pds_forward_f32 with a workspace, or a lower capacity_pct)",
                    need, (size_t)PDS_STATIC_SCRATCH_BYTES);
    void* ws = nullptr;
    if (need > 0 && (ws = device_scratch()) == nullptr) return PDS_ERR_HIP;
    return forward_impl(Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, out, ws,
                        need > 0 ? (size_t)PDS_STATIC_SCRATCH_BYTES : 0, false, stream);
}

int pds_forward_bins_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp, const float* radius,
                         int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius, int32_t mode,
                         const pds_tuning* t, float* out, void* bins, size_t bins_bytes, void* stream) {
    return forward_impl(Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, out, bins, bins_bytes, true,
                        stream);
}

int pds_backward_f32(const float* grad, const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                     const float* radius, int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                     int32_t mode, const pds_tuning* t, float* grad_Ep, float* grad_xp, float* grad_Mp,
                     void* workspace, size_t workspace_bytes, void* stream) {
    return backward_impl(grad, Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, grad_Ep, grad_xp,
                         grad_Mp, workspace, workspace_bytes, false, stream);
}

int pds_backward_bins_f32(const float* grad, const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                          const float* radius, int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                          int32_t mode, const pds_tuning* t, float* grad_Ep, float* grad_xp, float* grad_Mp,
                          void* bins, size_t bins_bytes, void* stream) {
    return backward_impl(grad, Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, grad_Ep, grad_xp,
                         grad_Mp, bins, bins_bytes, true, stream);
}

int pds_forward_bins_live_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                              const float* radius, int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                              int32_t mode, const pds_tuning* t, const int64_t* live_count, float* out, void* bins,
                              size_t bins_bytes, void* stream) {
    if (!live_count) return fail(PDS_ERR_ARGUMENT, "null live count pointer");
    return forward_impl(Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, out, bins, bins_bytes, true,
                        stream, live_count);
}

int pds_backward_bins_live_f32(const float* grad, const float* Ep, const float* xp, const float* Mp,
                               const int64_t* cp, const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                               int32_t max_pixel_radius, int32_t mode, const pds_tuning* t,
                               const int64_t* live_count, float* grad_Ep, float* grad_xp, float* grad_Mp, void* bins,
                               size_t bins_bytes, void* stream) {
    if (!live_count) return fail(PDS_ERR_ARGUMENT, "null live count pointer");
    return backward_impl(grad, Ep, xp, Mp, cp, radius, N, C, H, W, max_pixel_radius, mode, t, grad_Ep, grad_xp,
                         grad_Mp, bins, bins_bytes, true, stream, live_count);
}

int pds_bins_stats_get(const void* bins, size_t bins_bytes, int64_t N, int32_t C, int32_t H, int32_t W,
                       int32_t max_pixel_radius, int32_t mode, const pds_tuning* t, pds_bins_stats* out,
                       void* stream) {
    g_last_error[0] = '\0';
    if (!out) return fail(PDS_ERR_ARGUMENT, "null stats pointer");
    if (N < 0 || C <= 0 || H <= 0 || W <= 0) return fail(PDS_ERR_ARGUMENT, "invalid problem size");
    if (mode < PDS_MODE_AUTO || mode > PDS_MODE_DETERMINISTIC) return fail(PDS_ERR_ARGUMENT, "unknown mode %d", mode);
    const int rc = pds::binned_stats(bins, bins_bytes, N, image_args(C, H, W, max_pixel_radius), mode, t, out,
                                     (hipStream_t)stream);
    if (rc == PDS_ERR_HIP) return hip_fail(hipGetLastError(), "bins stats");
    if (rc) return fail(rc, "bins stats: %s", pds::binned_last_error());
    return PDS_OK;
}

}  // extern "C"
