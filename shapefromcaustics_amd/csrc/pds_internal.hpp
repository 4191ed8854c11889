// pds_internal.hpp -- argument blocks, the per-photon footprint box, and the
// wavefront-level "flattened" walk shared by the PDS kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprim/warp/warp_scan.hpp>
#include <stdint.h>

#include "pds_device.hpp"

namespace pds {

// Pixels per work item of the flattened walk (see walk_items).
constexpr int kFwdChunk = 2;
constexpr int kBwdChunk = 2;

// Photon inputs in the reference layout (PhotonDifferentialSplattig.cpp:19-22).
struct PhotonArgs {
    const float* E;        // [N]
    const float* xp;       // [2, N]
    const float* M;        // [2, 3, N]
    const int64_t* cp;     // [1, N]
    const float* radius;   // [N]
    int64_t N;
};

struct ImageArgs {
    int32_t C, H, W, cap;
    float fw, fh;          // (float)W, (float)H as in photon_differentials.cu:76-77
    int32_t ell_trivial;   // reference ellipse test (:93) is a no-op for every window offset
};

struct GradArgs {
    float* gE;   // [N]
    float* gx;   // [2, N]
    float* gM;   // [2, 3, N]
};

// Store the seven accumulated sums of photon i; gM[:,2] is always zero
// (photon_differentials.cu:210-218).
__device__ __forceinline__ void store_grads(const GradArgs& G, int64_t N, int64_t i, const float (&a)[7]) {
    G.gE[i] = a[0];
    G.gx[i] = a[1];
    G.gx[N + i] = a[2];
    G.gM[i] = a[3];
    G.gM[N + i] = a[4];
    G.gM[2 * N + i] = 0.0f;
    G.gM[3 * N + i] = a[5];
    G.gM[4 * N + i] = a[6];
    G.gM[5 * N + i] = 0.0f;
}

// Compiler-level ordering of LDS accesses between lanes of one wavefront.
// The LDS executes one wavefront's instructions in order, so preventing the
// compiler from reordering them is all intra-wave exchange needs.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int wave_inclusive_max(int v) {
    using WS = rocprim::warp_scan<int, 64>;
    __shared__ typename WS::storage_type st;  // unused by the DPP implementation
    int o;
    WS().inclusive_scan(v, o, st, rocprim::maximum<int>());
    return o;
}

__device__ __forceinline__ int64_t wave_inclusive_sum(int64_t v) {
    using WS = rocprim::warp_scan<long long, 64>;
    __shared__ typename WS::storage_type st;
    long long o;
    WS().inclusive_scan((long long)v, o, st, rocprim::plus<long long>());
    return (int64_t)o;
}

// One photon's footprint box, prepared for the flattened walk.
struct FlatRec {
    float cx, cy, e, inv_cpr;
    float m00, m01, m10, m11;
    int32_t x0, y0, cpr, bw;   // box origin (image pixels), chunks per row, box width
    int32_t plane;             // c * H * W
    int32_t nitems;            // rows * chunks per row (0: nothing to do)
    int64_t start;             // exclusive prefix of nitems over the wavefront
};

// Per-wavefront LDS copy of the 64 records of the current batch.
struct WaveRecs {
    float4 A[64];      // cx, cy, e, inv_cpr
    float4 B[64];      // m00, m01, m10, m11
    int4 Cq[64];       // x0, y0, cpr, bw
    int32_t plane[64];
    int64_t start[64];
    int32_t mark[64];  // record whose first item lands on this lane (-1: none)
};

// Footprint box of photon i: the reference window (photon_differentials.cu:80-96)
// clipped to the image, then -- for photons whose inputs are all finite -- to
// the bounding box of the ellipse |M d| < 1 widened by one pixel.  Pixels
// outside that ellipse have K = 0 and contribute exactly nothing, so the
// result is unchanged.  Non-finite photons keep the whole window so that NaN
// propagates into their gradients the way it does in the reference.
template <int CH>
__device__ __forceinline__ void setup_flat_rec(const PhotonArgs& P, const ImageArgs& I, int64_t i, bool forward,
                                               FlatRec& r) {
    r.nitems = 0;
    r.x0 = r.y0 = r.bw = 0;
    r.cpr = 1;
    r.inv_cpr = 1.0f;
    r.plane = 0;
    r.cx = r.cy = r.e = 0.f;
    r.m00 = r.m01 = r.m10 = r.m11 = 0.f;
    if (i >= P.N) return;
    const int64_t c = P.cp[i];
    if (c < 0 || c >= I.C) return;
    const int64_t N = P.N;
    const float cx = P.xp[i], cy = P.xp[N + i];
    const float e = P.E[i];
    if (forward && !(e > 0.0f)) return;  // val = K E > 0 never holds (:106)
    const Window win = photon_window(P.radius[i], cx, cy, I.W, I.H, I.cap);
    if (win.rx < 0 || win.ry < 0) return;
    int64_t x0 = max((int64_t)win.pxc - win.rx, (int64_t)0);
    int64_t x1 = min((int64_t)win.pxc + win.rx, (int64_t)I.W - 1);
    int64_t y0 = max((int64_t)win.pyc - win.ry, (int64_t)0);
    int64_t y1 = min((int64_t)win.pyc + win.ry, (int64_t)I.H - 1);
    if (x0 > x1 || y0 > y1) return;
    const float m00 = P.M[i], m01 = P.M[N + i], m10 = P.M[3 * N + i], m11 = P.M[4 * N + i];
    const float z = P.M[2 * N + i] * 0.0f + P.M[5 * N + i] * 0.0f;  // NaN iff M[:,2] is not finite
    const bool finite = isfinite(cx) && isfinite(cy) && isfinite(e) && isfinite(m00) && isfinite(m01) &&
                        isfinite(m10) && isfinite(m11) && z == 0.0f;
    if (finite) {
        const double a = m00, b = m01, cc = m10, d = m11;
        const double det = a * d - b * cc;
        if (det != 0.0) {
            const double ad = fabs(det);
            const double hx = sqrt(d * d + b * b) / ad * (0.5 * (double)I.W);
            const double hy = sqrt(cc * cc + a * a) / ad * (0.5 * (double)I.H);
            const double cxp = ((double)cx + 1.0) * 0.5 * (double)I.W;
            const double cyp = ((double)cy + 1.0) * 0.5 * (double)I.H;
            const double lx = fmin(fmax(floor(cxp - hx) - 1.0, (double)x0), (double)x1 + 1.0);
            const double hxp = fmax(fmin(ceil(cxp + hx) + 1.0, (double)x1), (double)x0 - 1.0);
            const double ly = fmin(fmax(floor(cyp - hy) - 1.0, (double)y0), (double)y1 + 1.0);
            const double hyp = fmax(fmin(ceil(cyp + hy) + 1.0, (double)y1), (double)y0 - 1.0);
            x0 = (int64_t)lx;
            x1 = (int64_t)hxp;
            y0 = (int64_t)ly;
            y1 = (int64_t)hyp;
            if (x0 > x1 || y0 > y1) return;
        }
    }
    const int32_t bw = (int32_t)(x1 - x0 + 1), bh = (int32_t)(y1 - y0 + 1);
    r.cpr = (bw + CH - 1) / CH;
    r.inv_cpr = 1.0f / (float)r.cpr;
    r.nitems = bh * r.cpr;
    r.x0 = (int32_t)x0;
    r.y0 = (int32_t)y0;
    r.bw = bw;
    r.plane = (int32_t)(c * I.H * I.W);
    // a NaN from M[:,2] enters through the centre so that u and v are NaN
    r.cx = finite ? cx : cx + z;
    r.cy = cy;
    r.e = e;
    r.m00 = m00;
    r.m01 = m01;
    r.m10 = m10;
    r.m11 = m11;
}

// Publish this lane's record to the wavefront's LDS table; returns the total
// number of work items of the batch and sets r.start.
__device__ __forceinline__ int64_t publish_recs(WaveRecs& R, FlatRec& r, int lane) {
    const int64_t incl = wave_inclusive_sum((int64_t)r.nitems);
    r.start = incl - r.nitems;
    const int64_t total = __shfl(incl, 63);
    R.A[lane] = make_float4(r.cx, r.cy, r.e, r.inv_cpr);
    R.B[lane] = make_float4(r.m00, r.m01, r.m10, r.m11);
    R.Cq[lane] = make_int4(r.x0, r.y0, r.cpr, r.bw);
    R.plane[lane] = r.plane;
    R.start[lane] = r.start;
    wave_lds_fence();
    return total;
}

// One work item: up to CH consecutive pixels of one row of one photon's box.
struct FlatItem {
    int32_t rec;             // lane (record) the item belongs to
    int32_t px, py;          // first pixel
    int32_t n;               // pixels in this item (1..CH)
    int64_t pix;             // flat index of the first pixel in the [C,H,W] image
    float cx, dy, e;
    float m00, m01, m10, m11;
};

// Walk all work items of the batch, 64 per step.  The lane -> record map of a
// step is an inclusive max-scan of "record starts here" marks: records are in
// lane order and their item ranges are contiguous, so the owner of item j is
// the last record that started at or before j.
template <int CH, class F>
__device__ __forceinline__ void walk_items(WaveRecs& R, const FlatRec& r, const ImageArgs& I, int lane, int64_t total,
                                           F&& body) {
    int carry = -1;
    for (int64_t base = 0; base < total; base += 64) {
        if (r.nitems > 0 && r.start >= base && r.start < base + 64) R.mark[(int)(r.start - base)] = lane;
        wave_lds_fence();
        int m = R.mark[lane];
        wave_lds_fence();
        R.mark[lane] = -1;
        if (lane == 0) m = max(m, carry);
        m = wave_inclusive_max(m);
        carry = __shfl(m, 63);
        const int64_t item = base + lane;
        if (item < total) {
            FlatItem it;
            it.rec = m;
            const float4 A = R.A[m];
            const float4 B = R.B[m];
            const int4 Q = R.Cq[m];
            const int32_t local = (int32_t)(item - R.start[m]);
            int32_t row = (int32_t)(((float)local + 0.5f) * A.w);
            if (local >= (1 << 20)) row = local / Q.z;  // the float quotient is exact below 2^20
            const int32_t chunk = local - row * Q.z;
            it.px = Q.x + chunk * CH;
            it.py = Q.y + row;
            it.n = min(CH, Q.w - chunk * CH);
            it.pix = (int64_t)R.plane[m] + (int64_t)it.py * I.W + it.px;
            it.cx = A.x;
            it.dy = pixel_coord(it.py, I.fh) - A.y;
            it.e = A.z;
            it.m00 = B.x;
            it.m01 = B.y;
            it.m10 = B.z;
            it.m11 = B.w;
            body(it);
        }
        wave_lds_fence();
    }
}

// launchers (pds_kernels.hip)
hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t s);
hipError_t launch_forward_plain(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s);
hipError_t launch_backward_plain(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                 hipStream_t s);
hipError_t launch_forward_flat(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s);
hipError_t launch_backward_flat(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                hipStream_t s);

}  // namespace pds
