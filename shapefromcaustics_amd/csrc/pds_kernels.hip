// pds_kernels.hip -- photon-differential splatting kernels for MI355X (gfx950).
//
// Two families:
//
//  * "plain": one thread per photon walking the reference window serially,
//    global FP32 atomics in the forward.  This is the reference's schedule
//    (PyOptix/kernel/photon_differentials.cu:60-113, :136-221), kept as the
//    plain-atomic comparator and for images small enough that the reference's
//    window ellipse test (:93) is not a no-op.
//
//  * "flat": each wavefront takes 64 photons, prunes each photon's window to
//    the bounding box of its elliptic footprint (exact: pixels outside it have
//    K = 0), and walks the concatenation of all 64 boxes in (photon, row,
//    CH-pixel chunk) work items, 64 items per step, so every lane does useful
//    work whatever the spread of footprint sizes.  The forward scatters with
//    global atomics or into an LDS tile (binned path, pds_binned.hip); the
//    backward gathers the upstream gradient and reduces the seven per-photon
//    sums in LDS, then stores them coalesced in photon order.
#include "pds_internal.hpp"

namespace pds {

// ---------------------------------------------------------------------------
// zero fill (a kernel, not hipMemsetAsync, so the forward can be graph-captured)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_zero_kernel(float4* __restrict__ p, int64_t n4, float* tail,
                                                        int32_t ntail) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (blockIdx.x == 0 && (int32_t)threadIdx.x < ntail) tail[threadIdx.x] = 0.f;
}

hipError_t launch_fill_zero(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    // float4 body from the first 16-byte aligned element, scalar head/tail
    int64_t head = (int64_t)(((16 - ((uintptr_t)p & 15)) & 15) / 4);
    if (head > n) head = n;
    int64_t body = (n - head) / 4;
    int64_t tail = n - head - body * 4;
    // head and tail are < 4 elements each; write them with one tiny launch
    if (head > 0) {
        hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, s, (float4*)nullptr, (int64_t)0, p, (int32_t)head);
    }
    if (body > 0 || tail > 0) {
        int64_t blocks = (body + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (float4*)(p + head), body,
                           p + head + body * 4, (int32_t)tail);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// plain kernels: the reference's schedule
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fwd_plain_kernel(PhotonArgs P, ImageArgs I, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    const PlainSetup s = plain_setup(P, I, i);
    if (!s.ok) return;
    float* plane = out + (int64_t)s.c * I.H * I.W;
    const Frame& f = s.f;
    for (int64_t py = s.y0; py <= s.y1; ++py) {
        const float dy = (float)((int32_t)py - f.iy) - f.fy;
        for (int64_t px = s.x0; px <= s.x1; ++px) {
            if (!I.ell_trivial && !window_ellipse_ok((int32_t)(px - s.win.pxc), (int32_t)(py - s.win.pyc), I.W, I.H))
                continue;
            const float dx = (float)((int32_t)px - f.ix) - f.fx;
            float u, v;
            const float val = silverman(eval_q(f, dx, dy, u, v)) * s.e;
            if (val > 0.0f) unsafeAtomicAdd(&plane[py * I.W + px], val);
        }
    }
}

__global__ __launch_bounds__(256) void bwd_plain_kernel(const float* __restrict__ grad, PhotonArgs P, ImageArgs I,
                                                        GradArgs G) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    float a[7];
    photon_grad_direct(grad, P, I, i, /*clip_box=*/false, a);
    store_grads(G, I, P.N, i, a);
}

// ---------------------------------------------------------------------------
// flat kernels (photon order)
// ---------------------------------------------------------------------------

// Forward: global FP32 atomics into the zero-filled image.
template <int U>
__global__ __launch_bounds__(256) void fwd_flat_kernel(PhotonArgs P, ImageArgs I, float* __restrict__ out) {
    __shared__ WaveRecs recs[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveRecs& R = recs[w];
    R.mark[lane] = -1;
    const int64_t nbatch = (P.N + 63) >> 6;
    for (int64_t b = (int64_t)blockIdx.x * 4 + w; b < nbatch; b += (int64_t)gridDim.x * 4) {
        const int64_t i = (b << 6) + lane;
        FlatRec r;
        setup_flat_rec<U>(P, I, i, /*forward=*/true, r, P.N);
        const int64_t total = publish_recs(R, r, lane);
        walk_units<U>(R, r, lane, total, [&](const FlatItem& it) { flat_forward_unit<U>(it, I.W, out); });
    }
}

// Backward: gather + per-photon LDS reduction, coalesced stores.  The LDS
// sums are double: gfx950's LDS float add serialises a wavefront's lanes
// (~190 cycles per instruction per CU), its double add does not (~9).
template <int U>
__global__ __launch_bounds__(256) void bwd_flat_kernel(const float* __restrict__ grad, PhotonArgs P, ImageArgs I,
                                                       GradArgs G) {
    __shared__ WaveRecs recs[4];
    __shared__ double acc[4][7][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveRecs& R = recs[w];
    R.mark[lane] = -1;
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[w][k][lane] = 0.0;
    const int64_t nbatch = (P.N + 63) >> 6;
    for (int64_t b = (int64_t)blockIdx.x * 4 + w; b < nbatch; b += (int64_t)gridDim.x * 4) {
        const int64_t i = (b << 6) + lane;
        FlatRec r;
        setup_flat_rec<U>(P, I, i, /*forward=*/false, r, P.N);
        const int64_t total = publish_recs(R, r, lane);
        walk_units<U>(R, r, lane, total, [&](const FlatItem& it) {
            const UnitPixels<U> px(it, I.W);
            float g[U];  // every gather of the unit is issued before any is used
#pragma unroll
            for (int k = 0; k < U; ++k) g[k] = grad[k < it.n ? px.pix[k] : px.pix[0]];
            float a[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < U; ++k) {
                if (k < it.n) {
                    float u, v;
                    const float q = eval_q(it.f, px.dx[k], px.dy[k], u, v);
                    const float sc = dk_over_r(q, it.e * g[k]);
                    accumulate_terms(a, q, g[k], sc, u, v, px.dx[k], px.dy[k], it.f);
                }
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) atomicAdd(&acc[w][k][it.rec], (double)a[k]);
        });
        wave_lds_fence();
        float s[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            s[k] = (float)acc[w][k][lane];
            acc[w][k][lane] = 0.0;
        }
        if (i < P.N) store_grads(G, I, P.N, i, s);
        wave_lds_fence();
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned flat_grid(int64_t N) {
    int64_t batches = (N + 63) / 64;
    int64_t blocks = (batches + 3) / 4;
    // enough resident waves to fill 256 CUs several times over; the kernels loop
    if (blocks > 8192) blocks = 8192;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

hipError_t launch_forward_plain(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    const int64_t blocks = (P.N + 255) / 256;
    hipLaunchKernelGGL(fwd_plain_kernel, dim3((unsigned)blocks), dim3(256), 0, s, P, I, out);
    return hipGetLastError();
}

hipError_t launch_backward_plain(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                 hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    const int64_t blocks = (P.N + 255) / 256;
    hipLaunchKernelGGL(bwd_plain_kernel, dim3((unsigned)blocks), dim3(256), 0, s, grad, P, I, G);
    return hipGetLastError();
}

hipError_t launch_forward_flat(const PhotonArgs& P, const ImageArgs& I, float* out, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    hipLaunchKernelGGL(fwd_flat_kernel<kFwdUnit>, dim3(flat_grid(P.N)), dim3(256), 0, s, P, I, out);
    return hipGetLastError();
}

hipError_t launch_backward_flat(const float* grad, const PhotonArgs& P, const ImageArgs& I, const GradArgs& G,
                                hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    hipLaunchKernelGGL(bwd_flat_kernel<kBwdUnit>, dim3(flat_grid(P.N)), dim3(256), 0, s, grad, P, I, G);
    return hipGetLastError();
}

}  // namespace pds
