// pds_device.hpp -- device-side helpers shared by every PDS kernel.
//
// Window semantics follow the reference kernels exactly
// (PyOptix/kernel/photon_differentials.cu):
//   :80-81  rx = min(int(ceil(radius * 0.5 * W)), cap)   (double arithmetic)
//   :46-51  pixel centre = int((0.5 * x + 0.5) * W)        (double, truncation)
//   :93     window ellipse test x_off^2/(W^2/4) + y_off^2/(H^2/4) <= 1 (double)
//   :96     pixel in [0, W) x [0, H)
//   :39-44  pixel coordinate 2 px / W - 1 (pixel corner, float)
//   :23-37  Silverman kernel K(q) = 3/pi (1 - q)^2 and its derivative
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pds {

constexpr float kThreeOverPi = 0.954929658551372f;     // 3 / pi   (fp32 of photon_differentials.cu:26)
constexpr float kTwelveOverPi = 3.819718634205488f;    // 12 / pi  (fp32 of photon_differentials.cu:34)

// float/double -> int32 with saturation and NaN -> 0, the conversion rule of
// the GPU instruction the reference compiles to.  Written out because an
// out-of-range C++ cast is undefined.
__device__ __forceinline__ int32_t sat_i32(double d) {
    if (!(d == d)) return 0;
    if (d >= 2147483647.0) return 2147483647;
    if (d <= -2147483648.0) return -2147483647 - 1;
    return (int32_t)d;
}

// Reference window of one photon (photon_differentials.cu:80-87).
struct Window {
    int32_t rx, ry, pxc, pyc;
};

__device__ __forceinline__ Window photon_window(float radius, float cx, float cy, int32_t W, int32_t H,
                                                int32_t cap) {
    const double w = (double)(float)W, h = (double)(float)H;
    Window win;
    int32_t a = sat_i32(ceil((double)radius * 0.5 * w));
    int32_t b = sat_i32(ceil((double)radius * 0.5 * h));
    win.rx = a < cap ? a : cap;
    win.ry = b < cap ? b : cap;
    win.pxc = sat_i32((0.5 * (double)cx + 0.5) * w);
    win.pyc = sat_i32((0.5 * (double)cy + 0.5) * h);
    return win;
}

// The reference's window ellipse test (photon_differentials.cu:93).  It only
// removes offsets for images smaller than about twice the window; the host
// decides per launch whether it can be skipped (see ellipse_trivial()).
__device__ __forceinline__ bool window_ellipse_ok(int32_t xo, int32_t yo, int32_t W, int32_t H) {
    const double w = (double)(float)W, h = (double)(float)H;
    const double a = (double)(float)((int64_t)xo * xo) / (0.25 * w * w);
    const double b = (double)(float)((int64_t)yo * yo) / (0.25 * h * h);
    return a + b <= 1.0;
}

// Photon-local pixel frame.  The reference evaluates the kernel at
// dx = (2 px / W - 1) - x in float (photon_differentials.cu:98-100), which for
// a non-power-of-two W rounds the pixel coordinate by up to 6e-8, a visible
// error for footprints a fraction of a pixel wide.  Here the photon centre in
// pixel units, W (x + 1) / 2, is split in double into an integer ix and a
// fraction fx in [-1/2, 1/2], so that
//     dx * W / 2 = (px - ix) - fx            (one float rounding),
// and the matrix is pre-scaled once per photon, M' = M diag(2/W, 2/H), so that
//     (u, v) = M' (px - ix - fx, py - iy - fy)  equals  M (dx, dy).
// The backward rescales its sums by W/2, H/2, 2/W, 2/H once per photon.
struct Frame {
    int32_t ix, iy;
    float fx, fy;
    float a00, a01, a10, a11;   // M'
};

__device__ __forceinline__ Frame photon_frame(float cx, float cy, float m00, float m01, float m10, float m11,
                                              float poison, int32_t W, int32_t H, float inv_hw, float inv_hh) {
    Frame f;
    const double hw = 0.5 * (double)W, hh = 0.5 * (double)H;
    const double px = ((double)cx + 1.0) * hw, py = ((double)cy + 1.0) * hh;  // exact in double
    // photons that reach this point have a non-empty window on the image, so a
    // finite centre lies within a few windows of it and fits an int32
    const double rx = (px == px && fabs(px) < 1.0e9) ? rint(px) : 0.0;
    const double ry = (py == py && fabs(py) < 1.0e9) ? rint(py) : 0.0;
    f.ix = (int32_t)rx;
    f.iy = (int32_t)ry;
    f.fx = (float)(px - rx) + poison;   // NaN/inf centre or NaN poison propagate
    f.fy = (float)(py - ry);
    f.a00 = m00 * inv_hw;               // inv_hw = 2 / W
    f.a01 = m01 * inv_hh;
    f.a10 = m10 * inv_hw;
    f.a11 = m11 * inv_hh;
    return f;
}

// Kahan's 2x2 determinant with FMA: a d - b c to about 1.5 ulp even under
// cancellation.
__device__ __forceinline__ float det2(float a, float b, float c, float d) {
    const float w = b * c;
    const float e = fmaf(-b, c, w);
    const float f = fmaf(a, d, -w);
    return f + e;
}

// Half extents, in pixels, of the ellipse |A d| < 1 (A = the frame matrix):
// the row norms of A^-1, widened by 1e-4 relative plus 1e-4 px, far more than
// the float error of the extents or of q near the boundary.  Returns false
// for a singular or non-finite matrix (the caller then keeps the window).
__device__ __forceinline__ bool ellipse_half_extents(const Frame& f, float& hx, float& hy) {
    const float det = fabsf(det2(f.a00, f.a01, f.a10, f.a11));
    if (!(det > 0.0f) || !(det < INFINITY)) return false;
    const float inv = __builtin_amdgcn_rcpf(det);
    hx = __builtin_amdgcn_sqrtf(fmaf(f.a11, f.a11, f.a01 * f.a01)) * inv * 1.0001f + 1e-4f;
    hy = __builtin_amdgcn_sqrtf(fmaf(f.a10, f.a10, f.a00 * f.a00)) * inv * 1.0001f + 1e-4f;
    if (!(hx < 1.0e6f)) hx = 1.0e6f;
    if (!(hy < 1.0e6f)) hy = 1.0e6f;
    return true;
}

// Clip [x0, x1] x [y0, y1] (absolute pixels) to the ellipse's bounding box.
__device__ __forceinline__ void clip_to_ellipse(const Frame& f, int32_t& x0, int32_t& x1, int32_t& y0,
                                                int32_t& y1) {
    float hx, hy;
    if (!ellipse_half_extents(f, hx, hy)) return;
    x0 = max(x0, f.ix + (int32_t)ceilf(f.fx - hx));
    x1 = min(x1, f.ix + (int32_t)floorf(f.fx + hx));
    y0 = max(y0, f.iy + (int32_t)ceilf(f.fy - hy));
    y1 = min(y1, f.iy + (int32_t)floorf(f.fy + hy));
}

// (u, v) = A (dx, dy) and q = u^2 + v^2 with the fused multiply-adds written
// out, so that every kernel (and every position of a pixel within a work
// unit) evaluates a pixel bit-identically; compiler contraction could
// otherwise differ between unrolled copies.
__device__ __forceinline__ float eval_q(const Frame& f, float dx, float dy, float& u, float& v) {
    u = fmaf(f.a00, dx, f.a01 * dy);
    v = fmaf(f.a10, dx, f.a11 * dy);
    return fmaf(u, u, v * v);
}

__device__ __forceinline__ float silverman(float q) {
    const float t = 1.0f - q;
    return q < 1.0f ? kThreeOverPi * t * t : 0.0f;
}

// dk / r  with dk = K'(r) E g (photon_differentials.cu:186-193): the factor
// that multiplies (u, v) in every position/matrix gradient term.  Finite at
// r = 0 (the reference divides 0/0 there); NaN when q is NaN, so a NaN
// propagates into all gradient terms exactly as the reference's u/r does.
__device__ __forceinline__ float dk_over_r(float q, float eg) {
    float s = q < 1.0f ? -kTwelveOverPi * (1.0f - q) : 0.0f;
    if (q != q) s = q;
    return s * eg;
}

}  // namespace pds
