/*
 * pds_cpu_baseline.c -- multithreaded (OpenMP) FP32 host-CPU implementation of
 * the photon-differential splatting formulae: the CPU baseline timed beside
 * the HIP product path.  TEST/BENCH INFRASTRUCTURE ONLY: loaded by tests/ and
 * bench.py's cpu_baseline leg, never by the product library.
 *
 * Same window semantics as the FP64 oracle (pds_oracle.c) and the reference
 * (PyOptix/kernel/photon_differentials.cu:60-113 forward, :136-221 backward):
 * double-precision window setup, pixel-corner coordinates, val > 0 filter in
 * the forward.  Forward: per-thread private images + parallel reduction.
 * Backward: parallel-for over photons (each photon owns its gradient slots).
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "pds_oracle_common.h"

static inline float silverman_f(float q) { return q < 1.0f ? (3.0f / (float)PDS_PI) * (1.0f - q) * (1.0f - q) : 0.0f; }

/* Per-thread private images of the forward, kept between calls (grown when
 * a larger problem comes) so that no call pays for allocating and faulting
 * in C x H x W floats per thread.  The baseline is called from one host
 * thread at a time. */
static float* g_priv = NULL;
static size_t g_priv_n = 0;

static float* private_images(size_t n)
{
    if (n > g_priv_n) {
        free(g_priv);
        g_priv = (float*)malloc(n * sizeof(float));
        g_priv_n = g_priv ? n : 0;
    }
    return g_priv;
}

int pds_cpu_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int pds_cpu_forward_f32(const float* E, const float* xp, const float* M, const int64_t* cp,
                        const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                        int32_t cap, float* out, int32_t nthreads)
{
    if (N < 0 || C <= 0 || H <= 0 || W <= 0) return -1;
    size_t img = (size_t)C * H * W;
    int nt = nthreads > 0 ? nthreads : pds_cpu_threads();
    const int ell_all = window_ellipse_ok(cap, cap, W, H);
    float* priv = private_images(img * (size_t)nt);
    if (!priv) return -2;
    const float fw = (float)W, fh = (float)H;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        float* mine = priv + img * (size_t)tid;
        memset(mine, 0, img * sizeof(float));  /* each thread clears its own image (first touch) */
#pragma omp for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            int64_t c = cp[i];
            if (c < 0 || c >= C) continue;
            int32_t rx, ry, pxc, pyc;
            float cx = xp[i], cy = xp[N + i];
            photon_window(radius[i], cx, cy, W, H, cap, &rx, &ry, &pxc, &pyc);
            if (rx < 0 || ry < 0) continue;
            float m00 = M[0 * N + i], m01 = M[1 * N + i], m10 = M[3 * N + i], m11 = M[4 * N + i];
            float e = E[i];
            float* plane = mine + (size_t)c * H * W;
            for (int64_t yo = -(int64_t)ry; yo <= ry; ++yo) {
                int64_t py = (int64_t)pyc + yo;
                if (py < 0 || py >= H) continue;
                float dy = 2.0f * (float)py / fh - 1.0f - cy;
                for (int64_t xo = -(int64_t)rx; xo <= rx; ++xo) {
                    if (!ell_all && !window_ellipse_ok(xo, yo, W, H)) continue;
                    int64_t px = (int64_t)pxc + xo;
                    if (px < 0 || px >= W) continue;
                    float dx = 2.0f * (float)px / fw - 1.0f - cx;
                    float u = m00 * dx + m01 * dy, v = m10 * dx + m11 * dy;
                    float val = silverman_f(u * u + v * v) * e;
                    if (val > 0) plane[(size_t)py * W + (size_t)px] += val;
                }
            }
        }
#pragma omp for schedule(static)
        for (int64_t p = 0; p < (int64_t)img; ++p) {
            float s = 0.0f;
            for (int t = 0; t < nt; ++t) s += priv[img * (size_t)t + (size_t)p];
            out[p] = s;
        }
    }
    return 0;
}

int pds_cpu_backward_f32(const float* grad, const float* E, const float* xp, const float* M,
                         const int64_t* cp, const float* radius, int64_t N, int32_t C, int32_t H,
                         int32_t W, int32_t cap, float* gE, float* gx, float* gM, int32_t nthreads)
{
    if (N < 0 || C <= 0 || H <= 0 || W <= 0) return -1;
    int nt = nthreads > 0 ? nthreads : pds_cpu_threads();
    const int ell_all = window_ellipse_ok(cap, cap, W, H);
    const float fw = (float)W, fh = (float)H;
    const float k12 = -12.0f / (float)PDS_PI;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < N; ++i) {
        float a[7] = {0, 0, 0, 0, 0, 0, 0};
        int64_t c = cp[i];
        if (c >= 0 && c < C) {
            int32_t rx, ry, pxc, pyc;
            float cx = xp[i], cy = xp[N + i];
            photon_window(radius[i], cx, cy, W, H, cap, &rx, &ry, &pxc, &pyc);
            float m00 = M[0 * N + i], m01 = M[1 * N + i], m10 = M[3 * N + i], m11 = M[4 * N + i];
            float e = E[i];
            const float* plane = grad + (size_t)c * H * W;
            if (rx >= 0 && ry >= 0) {
                for (int64_t yo = -(int64_t)ry; yo <= ry; ++yo) {
                    int64_t py = (int64_t)pyc + yo;
                    if (py < 0 || py >= H) continue;
                    float dy = 2.0f * (float)py / fh - 1.0f - cy;
                    for (int64_t xo = -(int64_t)rx; xo <= rx; ++xo) {
                        if (!ell_all && !window_ellipse_ok(xo, yo, W, H)) continue;
                        int64_t px = (int64_t)pxc + xo;
                        if (px < 0 || px >= W) continue;
                        float dx = 2.0f * (float)px / fw - 1.0f - cx;
                        float u = m00 * dx + m01 * dy, v = m10 * dx + m11 * dy;
                        float q = u * u + v * v;
                        float g = plane[(size_t)py * W + (size_t)px];
                        float s = (q < 1.0f ? k12 * (1.0f - q) : 0.0f) * e * g;
                        a[0] += silverman_f(q) * g;
                        a[1] -= s * (m00 * u + m10 * v);
                        a[2] -= s * (m01 * u + m11 * v);
                        a[3] += s * u * dx;
                        a[4] += s * u * dy;
                        a[5] += s * v * dx;
                        a[6] += s * v * dy;
                    }
                }
            }
        }
        gE[i] = a[0];
        gx[i] = a[1];
        gx[N + i] = a[2];
        gM[0 * N + i] = a[3];
        gM[1 * N + i] = a[4];
        gM[2 * N + i] = 0.0f;
        gM[3 * N + i] = a[5];
        gM[4 * N + i] = a[6];
        gM[5 * N + i] = 0.0f;
    }
    return 0;
}
