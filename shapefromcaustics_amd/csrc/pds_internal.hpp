// pds_internal.hpp -- argument blocks, the per-photon footprint box, and the
// wavefront-level "flattened" walk shared by the PDS kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprim/warp/warp_scan.hpp>
#include <stdint.h>

#include "pds_device.hpp"

namespace pds {

// Pixels per work item ("unit") of the flattened walk (see walk_units).
constexpr int kFwdUnit = 4;
constexpr int kBwdUnit = 8;

// Photon inputs in the reference layout (PhotonDifferentialSplattig.cpp:19-22).
struct PhotonArgs {
    const float* E;        // [N]
    const float* xp;       // [2, N]
    const float* M;        // [2, 3, N]
    const int64_t* cp;     // [1, N]
    const float* radius;   // [N]
    int64_t N;             // photons, and the stride of the [2, N] / [2, 3, N] planes
    const int64_t* live;   // null, or a device count: only photons i < *live take part (a batch's load)
};

// The photons that take part: [0, N), or [0, min(N, *live)) for a batch with
// a device-side count (read when the kernel runs, so that one captured graph
// serves any count up to N).
__device__ __forceinline__ int64_t live_photons(const PhotonArgs& P) {
    if (!P.live) return P.N;
    const int64_t n = *P.live;
    return n < 0 ? 0 : (n < P.N ? n : P.N);
}

struct ImageArgs {
    int32_t C, H, W, cap;
    float fw, fh;          // (float)W, (float)H as in photon_differentials.cu:76-77
    float inv_hw, inv_hh;  // 2 / W, 2 / H
    int32_t ell_trivial;   // reference ellipse test (:93) is a no-op for every window offset
};

struct GradArgs {
    float* gE;   // [N]
    float* gx;   // [2, N]
    float* gM;   // [2, 3, N]
};

// Accumulate the seven gradient terms of one pixel in the photon-local frame
// (photon_differentials.cu:195-204 with dk * (u, v) / r = sc * (u, v)):
//   a0 += K g;  a1 -= sc (a00 u + a10 v);  a2 -= sc (a01 u + a11 v);
//   a3 += sc u dx;  a4 += sc u dy;  a5 += sc v dx;  a6 += sc v dy
// where (dx, dy) are pixel offsets and a.. the pre-scaled matrix; store_grads
// converts back to the reference's normalised units.
__device__ __forceinline__ void accumulate_terms(float (&a)[7], float q, float g, float sc, float u, float v, float dx,
                                                 float dy, const Frame& f) {
    a[0] += silverman(q) * g;
    a[1] -= sc * (f.a00 * u + f.a10 * v);
    a[2] -= sc * (f.a01 * u + f.a11 * v);
    a[3] += sc * u * dx;
    a[4] += sc * u * dy;
    a[5] += sc * v * dx;
    a[6] += sc * v * dy;
}

// Store the seven sums of photon i in the reference's units and layout;
// gM[:,2] is always zero (photon_differentials.cu:210-218).  Pixel offsets
// are (W/2) dx and (H/2) dy, and a.. = M diag(2/W, 2/H), hence the factors.
__device__ __forceinline__ void store_grads(const GradArgs& G, const ImageArgs& I, int64_t N, int64_t i,
                                            const float (&a)[7]) {
    const float hw = 0.5f * I.fw, hh = 0.5f * I.fh;
    const float iw = 2.0f / I.fw, ih = 2.0f / I.fh;
    G.gE[i] = a[0];
    G.gx[i] = a[1] * hw;
    G.gx[N + i] = a[2] * hh;
    G.gM[i] = a[3] * iw;
    G.gM[N + i] = a[4] * ih;
    G.gM[2 * N + i] = 0.0f;
    G.gM[3 * N + i] = a[5] * iw;
    G.gM[4 * N + i] = a[6] * ih;
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

// One photon's footprint box, prepared for the flattened walk.  The box's
// pixels are numbered row-major; a work unit is U consecutive pixels of that
// order (it may wrap to the next row), so units carry the same work whatever
// the box shape and a photon needs ceil(area / U) of them.
struct FlatRec {
    Frame f;
    float e, inv_bw;
    int32_t kx0, ky0;          // box origin relative to (ix, iy)
    int32_t bw, area;          // box width, box area
    int32_t base;              // flat index of the box origin (image or LDS tile)
    int32_t nitems;            // units (0: nothing to do)
    int64_t start;             // exclusive prefix of nitems over the wavefront
};

__device__ __forceinline__ void clear_rec(FlatRec& r) {
    r.nitems = 0;
    r.kx0 = r.ky0 = r.base = r.area = 0;
    r.bw = 1;
    r.inv_bw = 1.0f;
    r.e = 0.f;
    r.f.ix = r.f.iy = 0;
    r.f.fx = r.f.fy = r.f.a00 = r.f.a01 = r.f.a10 = r.f.a11 = 0.f;
}

template <int U>
__device__ __forceinline__ void set_box(FlatRec& r, int32_t x0, int32_t y0, int32_t bw, int32_t bh, int32_t base) {
    r.bw = bw;
    r.inv_bw = 1.0f / (float)bw;
    r.area = bw * bh;
    r.nitems = (r.area + U - 1) / U;
    r.kx0 = x0 - r.f.ix;
    r.ky0 = y0 - r.f.iy;
    r.base = base;
}

// Per-wavefront LDS copy of the 64 records of the current batch.
struct WaveRecs {
    float4 A[64];      // fx, fy, e, inv_bw
    float4 B[64];      // a00, a01, a10, a11
    int4 Cq[64];       // kx0, ky0, bw, area
    int32_t base[64];
    int64_t start[64];
    int32_t mark[64];  // record whose first item lands on this lane (-1: none)
};

// Footprint box of photon i: the reference window (photon_differentials.cu:80-96)
// clipped to the image, then -- for photons whose inputs are all finite -- to
// the bounding box of the ellipse |M d| < 1 (with a 1e-4 relative safety
// margin).  Pixels outside that ellipse have K = 0 and contribute exactly
// nothing, so the result is unchanged.  Non-finite photons keep the whole
// window so that NaN propagates into their gradients as in the reference.
// n_end: photons i >= n_end are skipped (the live count of a batch)
template <int U>
__device__ __forceinline__ void setup_flat_rec(const PhotonArgs& P, const ImageArgs& I, int64_t i, bool forward,
                                               FlatRec& r, int64_t n_end) {
    clear_rec(r);
    if (i >= n_end) return;
    const int64_t c = P.cp[i];
    if (c < 0 || c >= I.C) return;
    const int64_t N = P.N;
    const float cx = P.xp[i], cy = P.xp[N + i];
    const float e = P.E[i];
    if (forward && !(e > 0.0f)) return;  // val = K E > 0 never holds (:106)
    const Window win = photon_window(P.radius[i], cx, cy, I.W, I.H, I.cap);
    if (win.rx < 0 || win.ry < 0) return;
    int32_t x0 = (int32_t)max((int64_t)win.pxc - win.rx, (int64_t)0);
    int32_t x1 = (int32_t)min((int64_t)win.pxc + win.rx, (int64_t)I.W - 1);
    int32_t y0 = (int32_t)max((int64_t)win.pyc - win.ry, (int64_t)0);
    int32_t y1 = (int32_t)min((int64_t)win.pyc + win.ry, (int64_t)I.H - 1);
    if (x0 > x1 || y0 > y1) return;
    const float m00 = P.M[i], m01 = P.M[N + i], m10 = P.M[3 * N + i], m11 = P.M[4 * N + i];
    const float poison = P.M[2 * N + i] * 0.0f + P.M[5 * N + i] * 0.0f;  // NaN iff M[:,2] is not finite
    const Frame f = photon_frame(cx, cy, m00, m01, m10, m11, poison, I.W, I.H, I.inv_hw, I.inv_hh);
    const bool finite = isfinite(cx) && isfinite(cy) && isfinite(e) && isfinite(m00) && isfinite(m01) &&
                        isfinite(m10) && isfinite(m11) && poison == 0.0f;
    if (finite) {
        clip_to_ellipse(f, x0, x1, y0, y1);
        if (x0 > x1 || y0 > y1) return;
    }
    r.f = f;
    r.e = e;
    set_box<U>(r, x0, y0, x1 - x0 + 1, y1 - y0 + 1, (int32_t)((c * I.H + y0) * I.W + x0));
}

// Publish this lane's record to the wavefront's LDS table; returns the total
// number of work items of the batch and sets r.start.
__device__ __forceinline__ int64_t publish_recs(WaveRecs& R, FlatRec& r, int lane) {
    const int64_t incl = wave_inclusive_sum((int64_t)r.nitems);
    r.start = incl - r.nitems;
    const int64_t total = __shfl(incl, 63);
    R.A[lane] = make_float4(r.f.fx, r.f.fy, r.e, r.inv_bw);
    R.B[lane] = make_float4(r.f.a00, r.f.a01, r.f.a10, r.f.a11);
    R.Cq[lane] = make_int4(r.kx0, r.ky0, r.bw, r.area);
    R.base[lane] = r.base;
    R.start[lane] = r.start;
    wave_lds_fence();
    return total;
}

// One work unit: up to U consecutive pixels (row-major) of one photon's box.
struct FlatItem {
    int32_t rec;              // lane (record) the unit belongs to
    int32_t row0, col0;       // first pixel, relative to the box origin
    int32_t n;                // pixels in this unit (1..U)
    int32_t kx0, ky0, bw;     // box origin relative to (ix, iy), box width
    int32_t base;             // flat index of the box origin
    float e;
    Frame f;                  // fx, fy, a.. valid (ix, iy unused)
};

// Walk all work units of the batch, 64 per step.  The lane -> record map of a
// step is an inclusive max-scan of "record starts here" marks: records are in
// lane order and their unit ranges are contiguous, so the owner of unit j is
// the last record that started at or before j.
template <int U, class F>
__device__ __forceinline__ void walk_units(WaveRecs& R, const FlatRec& r, int lane, int64_t total, F&& body) {
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
            const int32_t p0 = (int32_t)(item - R.start[m]) * U;
            // row = p0 / bw: the float estimate is within one of the quotient
            // for boxes of up to 2^18 pixels and is corrected exactly
            int32_t row = (int32_t)(((float)p0 + 0.5f) * A.w);
            if (p0 >= (1 << 18)) row = p0 / Q.z;
            int32_t col = p0 - row * Q.z;
            if (col < 0) {
                --row;
                col += Q.z;
            } else if (col >= Q.z) {
                ++row;
                col -= Q.z;
            }
            it.row0 = row;
            it.col0 = col;
            it.n = min(U, Q.w - p0);
            it.kx0 = Q.x;
            it.ky0 = Q.y;
            it.bw = Q.z;
            it.base = R.base[m];
            it.e = A.z;
            it.f.fx = A.x;
            it.f.fy = A.y;
            it.f.a00 = B.x;
            it.f.a01 = B.y;
            it.f.a10 = B.z;
            it.f.a11 = B.w;
            body(it);
        }
        wave_lds_fence();
    }
}

// Pixel k of a unit: offsets (dx, dy) from the photon centre in pixels and the
// flat index (row pitch ``pitch``).  Units are short, so the row/column of
// pixel k is stepped from the unit's first pixel.
template <int U>
struct UnitPixels {
    float dx[U], dy[U];
    int32_t pix[U];
    __device__ __forceinline__ UnitPixels(const FlatItem& it, int32_t pitch) {
        int32_t row = it.row0, col = it.col0;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            dx[k] = (float)(it.kx0 + col) - it.f.fx;
            dy[k] = (float)(it.ky0 + row) - it.f.fy;
            pix[k] = it.base + row * pitch + col;
            if (++col == it.bw) {
                col = 0;
                ++row;
            }
        }
    }
};

// Per-photon setup of the plain (reference-schedule) kernels: the reference
// window (photon_differentials.cu:74-89) clipped to the image, and the frame.
struct PlainSetup {
    bool ok;
    int32_t c;
    float e;
    Frame f;
    Window win;
    int64_t x0, x1, y0, y1;  // window clipped to the image (inclusive)
};

__device__ __forceinline__ PlainSetup plain_setup(const PhotonArgs& P, const ImageArgs& I, int64_t i) {
    PlainSetup s;
    s.ok = false;
    const int64_t c = P.cp[i];
    if (c < 0 || c >= I.C) return s;  // reference skips c >= C (:75); negative channels skipped too
    s.c = (int32_t)c;
    const float cx = P.xp[i], cy = P.xp[P.N + i];
    s.win = photon_window(P.radius[i], cx, cy, I.W, I.H, I.cap);
    if (s.win.rx < 0 || s.win.ry < 0) return s;
    s.x0 = max((int64_t)s.win.pxc - s.win.rx, (int64_t)0);
    s.x1 = min((int64_t)s.win.pxc + s.win.rx, (int64_t)I.W - 1);
    s.y0 = max((int64_t)s.win.pyc - s.win.ry, (int64_t)0);
    s.y1 = min((int64_t)s.win.pyc + s.win.ry, (int64_t)I.H - 1);
    s.ok = s.x0 <= s.x1 && s.y0 <= s.y1;
    if (!s.ok) return s;
    const int64_t N = P.N;
    // the reference multiplies M[:,2] by the zero third coordinate (:103): NaN if not finite
    const float poison = P.M[2 * N + i] * 0.0f + P.M[5 * N + i] * 0.0f;
    s.f = photon_frame(cx, cy, P.M[i], P.M[N + i], P.M[3 * N + i], P.M[4 * N + i], poison, I.W, I.H, I.inv_hw,
                       I.inv_hh);
    s.e = P.E[i];
    return s;
}

// The seven gradient sums of photon i by a serial walk of its window (the
// reference's schedule, photon_differentials.cu:172-208).  With clip_box the
// walk of a photon with finite inputs is limited to its footprint box; the
// result is the same.
__device__ __forceinline__ void photon_grad_direct(const float* __restrict__ grad, const PhotonArgs& P,
                                                   const ImageArgs& I, int64_t i, bool clip_box, float (&a)[7]) {
#pragma unroll
    for (int k = 0; k < 7; ++k) a[k] = 0.f;
    PlainSetup s = plain_setup(P, I, i);
    if (!s.ok) return;
    const Frame& f = s.f;
    if (clip_box) {
        const int64_t N = P.N;
        const bool finite = isfinite(f.fx) && isfinite(f.fy) && isfinite(s.e) && isfinite(P.M[i]) &&
                            isfinite(P.M[N + i]) && isfinite(P.M[3 * N + i]) && isfinite(P.M[4 * N + i]);
        if (finite) {
            int32_t x0 = (int32_t)s.x0, x1 = (int32_t)s.x1, y0 = (int32_t)s.y0, y1 = (int32_t)s.y1;
            clip_to_ellipse(f, x0, x1, y0, y1);
            s.x0 = x0;
            s.x1 = x1;
            s.y0 = y0;
            s.y1 = y1;
        }
    }
    const float* plane = grad + (int64_t)s.c * I.H * I.W;
    for (int64_t py = s.y0; py <= s.y1; ++py) {
        const float dy = (float)((int32_t)py - f.iy) - f.fy;
        for (int64_t px = s.x0; px <= s.x1; ++px) {
            if (!I.ell_trivial && !window_ellipse_ok((int32_t)(px - s.win.pxc), (int32_t)(py - s.win.pyc), I.W, I.H))
                continue;
            const float dx = (float)((int32_t)px - f.ix) - f.fx;
            float u, v;
            const float q = eval_q(f, dx, dy, u, v);
            const float g = plane[py * I.W + px];
            const float sc = dk_over_r(q, s.e * g);
            accumulate_terms(a, q, g, sc, u, v, dx, dy, f);
        }
    }
}

// launchers (pds_kernels.hip)
hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t s);
hipError_t launch_forward_plain(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s);
hipError_t launch_backward_plain(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                 hipStream_t s);
hipError_t launch_forward_flat(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s);
// The flat forward's body (global float atomics), shared with the binned
// module's capacity-overflow path.
template <int U>
__device__ __forceinline__ void flat_forward_unit(const FlatItem& it, int32_t W, float* __restrict__ out) {
    const UnitPixels<U> px(it, W);
#pragma unroll
    for (int k = 0; k < U; ++k) {
        if (k < it.n) {
            float u, v;
            const float val = silverman(eval_q(it.f, px.dx[k], px.dy[k], u, v)) * it.e;
            if (val > 0.0f) unsafeAtomicAdd(&out[px.pix[k]], val);
        }
    }
}
hipError_t launch_backward_flat(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                hipStream_t s);

}  // namespace pds
