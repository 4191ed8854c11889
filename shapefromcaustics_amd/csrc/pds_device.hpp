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

// Pixel-corner coordinate of pixel p along an axis of n pixels, in float with
// the reference's operation order (2 * p / n - 1, IEEE division).
__device__ __forceinline__ float pixel_coord(int32_t p, float n) { return 2.0f * (float)p / n - 1.0f; }

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
