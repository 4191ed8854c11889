/*
 * pds_oracle_common.h -- window helpers shared by the FP64 oracle and the
 * OpenMP CPU baseline.  TEST INFRASTRUCTURE ONLY (see pds_oracle.c).
 */
#ifndef PDS_ORACLE_COMMON_H
#define PDS_ORACLE_COMMON_H
#include <math.h>
#include <stdint.h>

#define PDS_PI 3.14159265358979323846

/* Saturating double -> int32 conversion, NaN -> 0 (GPU cvt semantics). */
static inline int32_t sat_i32(double d)
{
    if (!(d == d)) return 0;
    if (d >= 2147483647.0) return 2147483647;
    if (d <= -2147483648.0) return (int32_t)(-2147483647 - 1);
    return (int32_t)d; /* truncation toward zero */
}

/* Window of photon i, reference photon_differentials.cu:80-87 (double math). */
static inline void photon_window(float radius, float cx, float cy, int32_t W, int32_t H, int32_t cap,
                          int32_t* rx, int32_t* ry, int32_t* pxc, int32_t* pyc)
{
    double w = (double)(float)W, h = (double)(float)H;
    int32_t a = sat_i32(ceil((double)radius * 0.5 * w));
    int32_t b = sat_i32(ceil((double)radius * 0.5 * h));
    *rx = a < cap ? a : cap;
    *ry = b < cap ? b : cap;
    *pxc = sat_i32((0.5 * (double)cx + 0.5) * w);
    *pyc = sat_i32((0.5 * (double)cy + 0.5) * h);
}

/* The reference's (no-op for large images) ellipse test on window offsets,
 * photon_differentials.cu:93, in double. */
static inline int window_ellipse_ok(int64_t xo, int64_t yo, int32_t W, int32_t H)
{
    double w = (double)(float)W, h = (double)(float)H;
    double a = (double)(float)(xo * xo) / (0.25 * w * w);
    double b = (double)(float)(yo * yo) / (0.25 * h * h);
    return a + b <= 1.0;
}

static inline double silverman_d(double q) { return q < 1.0 ? 3.0 / PDS_PI * (1.0 - q) * (1.0 - q) : 0.0; }

#endif
