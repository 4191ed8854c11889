/*
 * pds_oracle.c -- CPU restatement of photon-differential splatting (PDS).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path and the CPU baseline timed beside it; it is never linked into, called
 * by, or used as a fallback for the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py (cpu_baseline leg) load it.
 *
 * Functions:
 *
 *   pds_oracle_forward_f64 / pds_oracle_backward_f64
 *       Single-threaded FP64 restatement of the reference kernels, written
 *       as the obviously-correct window loop of the reference (no pruning).
 *       Inputs are the same float32/int64 arrays the GPU sees; arithmetic is
 *       double.  The backward also returns, per output component, the sum of
 *       absolute values of its terms (the scale used by the gradient
 *       tolerance, since gx/gM cancel heavily on smooth upstream gradients).
 *
 *   (The OpenMP FP32 host baseline of the same formulae lives in
 *   pds_cpu_baseline.c; both share the window helpers of pds_oracle_common.h.)
 *
 * Reference semantics followed (paths relative to the reference repository):
 *   PyOptix/kernel/photon_differentials.cu:23-29   silverman kernel K(q)
 *   PyOptix/kernel/photon_differentials.cu:31-37   d_silverman K'(r)
 *   PyOptix/kernel/photon_differentials.cu:39-44   pixel_to_coord (corner, no +0.5)
 *   PyOptix/kernel/photon_differentials.cu:46-51   coord_to_pixel (double, trunc)
 *   PyOptix/kernel/photon_differentials.cu:53-58   matrix_multiply (M column 2 x 0)
 *   PyOptix/kernel/photon_differentials.cu:60-113  forward kernel (window, val>0)
 *   PyOptix/kernel/photon_differentials.cu:136-221 backward kernel (gather)
 *
 * Deliberate, documented deviations (shared with the HIP product path):
 *   - photons with channel < 0 are skipped (the reference indexes out of
 *     bounds); skipped photons get zero gradients (the reference leaves
 *     empty_like memory uninitialised);
 *   - at r = 0 the reference divides 0/0; here the gx/gM terms take their
 *     limit, 0 (the GPU path computes dk/r = -(12/pi)(1-q)E g directly);
 *   - float->int conversions saturate (NaN -> 0), which is what the GPU
 *     conversion instruction of the reference's target does;
 *   - a window with rx < 0 or ry < 0 is empty.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pds_oracle_common.h"

/* Forward, FP64.  out: C*H*W doubles, zero-filled here. */
int pds_oracle_forward_f64(const float* E, const float* xp, const float* M, const int64_t* cp,
                           const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                           int32_t cap, double* out)
{
    if (N < 0 || C <= 0 || H <= 0 || W <= 0) return -1;
    memset(out, 0, sizeof(double) * (size_t)C * (size_t)H * (size_t)W);
    for (int64_t i = 0; i < N; ++i) {
        int64_t c = cp[i];
        if (c < 0 || c >= C) continue;
        int32_t rx, ry, pxc, pyc;
        float cx = xp[i], cy = xp[N + i];
        photon_window(radius[i], cx, cy, W, H, cap, &rx, &ry, &pxc, &pyc);
        if (rx < 0 || ry < 0) continue;
        double m00 = M[0 * N + i], m01 = M[1 * N + i], m02 = M[2 * N + i];
        double m10 = M[3 * N + i], m11 = M[4 * N + i], m12 = M[5 * N + i];
        double e = E[i];
        for (int64_t yo = -(int64_t)ry; yo <= ry; ++yo) {
            for (int64_t xo = -(int64_t)rx; xo <= rx; ++xo) {
                if (!window_ellipse_ok(xo, yo, W, H)) continue;
                int64_t px = (int64_t)pxc + xo, py = (int64_t)pyc + yo;
                if (px < 0 || py < 0 || px >= W || py >= H) continue;
                double dx = 2.0 * (double)px / (double)W - 1.0 - (double)cx;
                double dy = 2.0 * (double)py / (double)H - 1.0 - (double)cy;
                double u = m00 * dx + m01 * dy + m02 * 0.0;
                double v = m10 * dx + m11 * dy + m12 * 0.0;
                double val = silverman_d(u * u + v * v) * e;
                if (val > 0) out[((size_t)c * H + (size_t)py) * W + (size_t)px] += val;
            }
        }
    }
    return 0;
}

/* Backward, FP64, over the photon subset idx[0..n_idx) (idx == NULL: all N).
 * Outputs are indexed by position in the subset: gE[k], gx[2][n], gM[6][n]
 * with n = n_idx.  a* (may be NULL) receive sums of |term| per component and
 * c* (may be NULL) a first-order bound of the FP32 evaluation error of each
 * component (conditioning of the photon's gradient). */
int pds_oracle_backward_f64(const float* grad, const float* E, const float* xp, const float* M,
                            const int64_t* cp, const float* radius, int64_t N, int32_t C,
                            int32_t H, int32_t W, int32_t cap, const int64_t* idx, int64_t n_idx,
                            double* gE, double* gx, double* gM, double* aE, double* ax, double* aM,
                            double* cE, double* cx_, double* cM)
{
    if (N < 0 || C <= 0 || H <= 0 || W <= 0) return -1;
    int64_t n = idx ? n_idx : N;
    for (int64_t k = 0; k < n; ++k) {
        int64_t i = idx ? idx[k] : k;
        double acc[9] = {0}, mag[9] = {0}, cond[9] = {0};
        int64_t c = cp[i];
        if (c >= 0 && c < C) {
            int32_t rx, ry, pxc, pyc;
            float cx = xp[i], cy = xp[N + i];
            photon_window(radius[i], cx, cy, W, H, cap, &rx, &ry, &pxc, &pyc);
            double m00 = M[0 * N + i], m01 = M[1 * N + i], m02 = M[2 * N + i];
            double m10 = M[3 * N + i], m11 = M[4 * N + i], m12 = M[5 * N + i];
            double e = E[i];
            if (rx >= 0 && ry >= 0) {
                for (int64_t yo = -(int64_t)ry; yo <= ry; ++yo) {
                    for (int64_t xo = -(int64_t)rx; xo <= rx; ++xo) {
                        if (!window_ellipse_ok(xo, yo, W, H)) continue;
                        int64_t px = (int64_t)pxc + xo, py = (int64_t)pyc + yo;
                        if (px < 0 || py < 0 || px >= W || py >= H) continue;
                        double dx = 2.0 * (double)px / (double)W - 1.0 - (double)cx;
                        double dy = 2.0 * (double)py / (double)H - 1.0 - (double)cy;
                        double u = m00 * dx + m01 * dy + m02 * 0.0;
                        double v = m10 * dx + m11 * dy + m12 * 0.0;
                        double q = u * u + v * v;
                        double r = sqrt(q);
                        double g = grad[((size_t)c * H + (size_t)py) * W + (size_t)px];
                        /* dk = K'(r) E g, photon_differentials.cu:186-193 */
                        double dk = (r < 1.0 ? -12.0 / PDS_PI * r * (1.0 - q) : 0.0) * e * g;
                        double t[9];
                        t[0] = silverman_d(q) * g;
                        if (r == 0.0) {
                            /* documented deviation: the limit r -> 0 of dk*(u/r) is 0
                             * (the reference evaluates 0/0 here) */
                            for (int j = 1; j < 9; ++j) t[j] = 0.0;
                        } else {
                            double un = u / r, vn = v / r; /* :189-190 */
                            t[1] = -dk * (m00 * un + m10 * vn); /* :196 */
                            t[2] = -dk * (m01 * un + m11 * vn); /* :197 */
                            t[3] = dk * un * dx;                /* :200 */
                            t[4] = dk * un * dy;                /* :201 */
                            t[5] = 0.0;                         /* :202, :215 */
                            t[6] = dk * vn * dx;                /* :203 */
                            t[7] = dk * vn * dy;                /* :204 */
                            t[8] = 0.0;                         /* :205, :218 */
                        }
                        for (int j = 0; j < 9; ++j) {
                            acc[j] += t[j];
                            mag[j] += fabs(t[j]);
                        }
                        /* First-order bound of what FP32 evaluation of this pixel's
                         * terms can be off by: rounding of u, v (a few ulps of
                         * their products, plus the photon centre held to 2^-25 px)
                         * and of q, propagated through each term's derivative.
                         * Large only where a footprint barely contains a pixel
                         * (1 - q tiny), where the result is ill-conditioned. */
                        if (q < 1.0 && r > 0.0) {
                            const double ur = 8.0 * 5.9604644775390625e-08; /* 8 * 2^-24 */
                            const double ddx = 2.9802322387695312e-08 * 2.0 / (double)W;
                            const double ddy = 2.9802322387695312e-08 * 2.0 / (double)H;
                            double eu = ur * (fabs(m00 * dx) + fabs(m01 * dy)) + fabs(m00) * ddx + fabs(m01) * ddy;
                            double ev = ur * (fabs(m10 * dx) + fabs(m11 * dy)) + fabs(m10) * ddx + fabs(m11) * ddy;
                            double eq = 2.0 * fabs(u) * eu + 2.0 * fabs(v) * ev + ur * q;
                            double sv = -12.0 / PDS_PI * (1.0 - q) * e * g; /* dk / r */
                            double sq = 12.0 / PDS_PI * e * g;               /* d(sv)/dq */
                            double gu = m00 * u + m10 * v, gv = m01 * u + m11 * v;
                            double d[9];
                            d[0] = 6.0 / PDS_PI * (1.0 - q) * fabs(g) * eq;
                            d[1] = fabs(sq * gu) * eq + fabs(sv) * (fabs(m00) * eu + fabs(m10) * ev);
                            d[2] = fabs(sq * gv) * eq + fabs(sv) * (fabs(m01) * eu + fabs(m11) * ev);
                            d[3] = fabs(sq * u * dx) * eq + fabs(sv * dx) * eu + fabs(sv * u) * ddx;
                            d[4] = fabs(sq * u * dy) * eq + fabs(sv * dy) * eu + fabs(sv * u) * ddy;
                            d[5] = 0.0;
                            d[6] = fabs(sq * v * dx) * eq + fabs(sv * dx) * ev + fabs(sv * v) * ddx;
                            d[7] = fabs(sq * v * dy) * eq + fabs(sv * dy) * ev + fabs(sv * v) * ddy;
                            d[8] = 0.0;
                            for (int j = 0; j < 9; ++j) cond[j] += d[j];
                        }
                    }
                }
            }
        }
        gE[k] = acc[0];
        gx[k] = acc[1];
        gx[n + k] = acc[2];
        for (int j = 0; j < 6; ++j) gM[(size_t)j * n + k] = acc[3 + j];
        if (aE) aE[k] = mag[0];
        if (ax) {
            ax[k] = mag[1];
            ax[n + k] = mag[2];
        }
        if (aM)
            for (int j = 0; j < 6; ++j) aM[(size_t)j * n + k] = mag[3 + j];
        if (cE) cE[k] = cond[0];
        if (cx_) {
            cx_[k] = cond[1];
            cx_[n + k] = cond[2];
        }
        if (cM)
            for (int j = 0; j < 6; ++j) cM[(size_t)j * n + k] = cond[3 + j];
    }
    return 0;
}


/* Loss of one photon against an upstream gradient image, all in double:
 *   L = sum over the photon's window of g[c,py,px] * K(|M (p - x)|^2) * E,
 * i.e. <g, F_i> with F_i the photon's forward contribution (the val > 0
 * filter is irrelevant for E > 0).  Parameters are doubles so that central
 * finite differences of L with respect to x and M can check the adjoint. */
double pds_oracle_photon_loss_f64(const float* grad, int32_t C, int32_t H, int32_t W, int32_t cap, int64_t c,
                                  double E, double x0, double x1, const double* M6, double radius)
{
    if (c < 0 || c >= C) return 0.0;
    double w = (double)(float)W, h = (double)(float)H;
    int32_t a = sat_i32(ceil(radius * 0.5 * w)), b = sat_i32(ceil(radius * 0.5 * h));
    int32_t rx = a < cap ? a : cap, ry = b < cap ? b : cap;
    if (rx < 0 || ry < 0) return 0.0;
    int32_t pxc = sat_i32((0.5 * x0 + 0.5) * w), pyc = sat_i32((0.5 * x1 + 0.5) * h);
    double L = 0.0;
    for (int64_t yo = -(int64_t)ry; yo <= ry; ++yo) {
        for (int64_t xo = -(int64_t)rx; xo <= rx; ++xo) {
            if (!window_ellipse_ok(xo, yo, W, H)) continue;
            int64_t px = (int64_t)pxc + xo, py = (int64_t)pyc + yo;
            if (px < 0 || py < 0 || px >= W || py >= H) continue;
            double dx = 2.0 * (double)px / w - 1.0 - x0;
            double dy = 2.0 * (double)py / h - 1.0 - x1;
            double u = M6[0] * dx + M6[1] * dy, v = M6[3] * dx + M6[4] * dy;
            L += grad[((size_t)c * H + (size_t)py) * W + (size_t)px] * silverman_d(u * u + v * v) * E;
        }
    }
    return L;
}
