// gen_kernels.hip -- counter-based generator of caustic-line photons
// (include/pds_gen.h).
//
// Photon i of view v is a pure function of (seed, v, i): every random draw is
// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) applied to a 64-bit counter built from the key
// and (i, draw), so a backward pass can regenerate the photons of a forward
// pass bit for bit instead of keeping them in memory.
//
// Distribution per view (the caustic-like splat): with probability 0.8 the
// photon lies on one of n_lines thin Gaussian lines (centre, axis angle and
// long sigma per line given by the caller, short sigma = long / 20), else it
// is uniform on the image; pixel-space footprint axes s_major ~ logU(0.3, 6)
// px, condition kappa ~ U(1, 10), orientation U(0, pi), energy U(0.5, 1.5)/N
// (N = photons of the view), channel = 3 v + i mod 3.  The mapping to the
// splat inputs is the one of model/caustics.py:224-239 for a footprint
// D = R(theta) diag(s1, s2): M[:, :2] = D^-1 diag(W/2, H/2), M[:, 2] = 0,
// radius = 2 max(s1, s2) / W, xp = 2 p / W - 1.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "pds.h"
#include "pds_gen.h"

namespace gen {

namespace {

thread_local char g_err[256] = "";

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1) with 24 random bits (exact in float)
__device__ __forceinline__ float u24(uint64_t bits) { return (float)(uint32_t)(bits >> 40) * (1.0f / 16777216.0f); }

struct GenArgs {
    uint64_t key;  // seed and view
    int64_t n;
    int32_t W, H, view, n_lines;
    const float* lines;  // [n_lines, 4]: mu_x, mu_y (px), axis angle, sigma_long (px)
    float inv_n, two_inv_w, two_inv_h;  // 1/n, 2/W, 2/H
    float *E, *xp, *M, *radius;
    int64_t* cp;
    int64_t stride;  // N of the destination arrays (xp [2, stride], M [2, 3, stride])
    int64_t offset;  // first destination index
};

__global__ __launch_bounds__(256) void caustic_kernel(GenArgs A) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.n) return;
    // draw k of photon i: splitmix64(key ^ (i << 4 | k)) -- distinct counters for k < 16
    const uint64_t ctr = (uint64_t)i << 4;
    const uint64_t r0 = splitmix64(A.key ^ (ctr | 0)), r1 = splitmix64(A.key ^ (ctr | 1));
    const uint64_t r2 = splitmix64(A.key ^ (ctr | 2)), r3 = splitmix64(A.key ^ (ctr | 3));
    const uint64_t r4 = splitmix64(A.key ^ (ctr | 4));
    const float u_line = u24(r0), u_k = u24(r0 << 24);
    const float u1 = u24(r1), u2 = u24(r1 << 24);       // Box-Muller
    const float ux = u24(r2), uy = u24(r2 << 24);       // uniform position
    const float u_th = u24(r3), u_s = u24(r3 << 24);
    const float u_kap = u24(r4), u_e = u24(r4 << 24);
    const float W = (float)A.W, H = (float)A.H;
    float px, py;
    if (u_line >= 0.2f) {
        const int k = min((int)(u_k * (float)A.n_lines), A.n_lines - 1);
        const float* L = A.lines + 4 * k;
        const float rad = sqrtf(-2.0f * logf(1.0f - u1));  // 1 - u1 in (0, 1]
        float sb, cb, c, s;
        sincosf(6.28318530717958648f * u2, &sb, &cb);  // one range reduction for both
        sincosf(L[2], &s, &c);
        const float z1 = rad * cb, z2 = rad * sb;
        const float a = L[3] * z1, b = (L[3] * 0.05f) * z2;
        px = L[0] + c * a - s * b;
        py = L[1] + s * a + c * b;
    } else {
        px = ux * W;
        py = uy * H;
    }
    px = fminf(fmaxf(px, 0.0f), W - 1e-3f);
    py = fminf(fmaxf(py, 0.0f), H - 1e-3f);
    const float theta = 3.14159265358979324f * u_th;
    const float s1 = 0.3f * expf(u_s * logf(20.0f));  // logU(0.3, 6)
    const float kappa = 1.0f + 9.0f * u_kap;
    float c, s;
    sincosf(theta, &s, &c);
    // reciprocals instead of divisions (the kernel is bound by its VALU
    // instructions): s2 = s1 / kappa, so 1/s2 = kappa/s1 and max(s1, s2) = s1
    // (kappa >= 1); 1/n, 2/W and 2/H come from the host
    const float inv_s1 = 1.0f / s1, inv_s2 = kappa * inv_s1;
    const int64_t j = A.offset + i, N = A.stride;
    A.E[j] = (0.5f + u_e) * A.inv_n;
    A.xp[j] = px * A.two_inv_w - 1.0f;
    A.xp[N + j] = py * A.two_inv_h - 1.0f;
    // D^-1 = diag(1/s1, 1/s2) R^T, times diag(W/2, H/2)
    A.M[j] = c * inv_s1 * (0.5f * W);
    A.M[N + j] = s * inv_s1 * (0.5f * H);
    A.M[2 * N + j] = 0.0f;
    A.M[3 * N + j] = -s * inv_s2 * (0.5f * W);
    A.M[4 * N + j] = c * inv_s2 * (0.5f * H);
    A.M[5 * N + j] = 0.0f;
    A.radius[j] = s1 * A.two_inv_w;
    A.cp[j] = 3 * (int64_t)A.view + (i % 3);
}

}  // namespace

}  // namespace gen

using namespace gen;

extern "C" {

const char* pds_gen_last_error(void) { return g_err; }

int pds_gen_caustic_f32(uint64_t seed, int32_t view, int64_t n, int32_t W, int32_t H, const float* lines,
                        int32_t n_lines, float* Ep, float* xp, float* Mp, int64_t* cp, float* radius,
                        int64_t stride, int64_t offset, void* stream) {
    g_err[0] = '\0';
    if (n < 0 || W <= 0 || H <= 0 || view < 0 || n_lines < 1 || !lines || stride < 0 || offset < 0 ||
        offset + n > stride) {
        snprintf(g_err, sizeof(g_err), "invalid generator arguments");
        return PDS_ERR_ARGUMENT;
    }
    if (n == 0) return PDS_OK;
    if (!Ep || !xp || !Mp || !cp || !radius) {
        snprintf(g_err, sizeof(g_err), "null output buffer");
        return PDS_ERR_ARGUMENT;
    }
    GenArgs A;
    A.key = (seed * 0x100000001B3ull) ^ ((uint64_t)(uint32_t)view << 56);
    A.n = n;
    A.W = W;
    A.H = H;
    A.view = view;
    A.n_lines = n_lines;
    A.lines = lines;
    A.inv_n = (float)(1.0 / (double)n);
    A.two_inv_w = (float)(2.0 / (double)W);
    A.two_inv_h = (float)(2.0 / (double)H);
    A.E = Ep;
    A.xp = xp;
    A.M = Mp;
    A.radius = radius;
    A.cp = cp;
    A.stride = stride;
    A.offset = offset;
    hipLaunchKernelGGL(caustic_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, A);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "generator launch: %s", hipGetErrorString(e));
        return PDS_ERR_HIP;
    }
    return PDS_OK;
}

}  // extern "C"
