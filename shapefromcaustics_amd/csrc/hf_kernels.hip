// hf_kernels.hip -- height-field caustic kernels (include/hf.h): Sobel
// normals, the trace contract, the fused photon path (emit -> trace ->
// transfer -> refract -> ... -> ground -> splat inputs), its reverse-mode
// adjoint with recomputation, and the Landweber data term and update.
//
// One thread per photon.  The path kernels keep a photon's whole path in
// registers; the adjoint re-runs it, keeping a per-bounce tape (20 floats a
// bounce, at most HF_MAX_BOUNCES bounces) in private memory, and sweeps it
// backwards.
// Per-vertex gradients (Sobel normals, vertex heights) gather per workgroup
// in an LDS patch (a 16x16 tile of the ray grid lands on a few cells), are
// flushed into [R, R] arrays with device float atomics, and one gather kernel
// applies the transposed Sobel/normalisation to give dL/dh.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "hf.h"
#include "hf_device.hpp"
#include "pds.h"

namespace hf {

namespace {

thread_local char g_err[256] = "";

int fail(int code, const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// workspace: [0] hmax (float) ; then grad_n [R,R,3], grad_z [R,R], grad_g [R,R,2]
constexpr int kHmaxBlocks = 128;
// path counters (hf_hits_stats_get): reached, missed, degenerate, capped
constexpr int kStatWords = 4;

struct WsLayout {
    size_t hmax, grad_n, grad_z, grad_g, end;
};
WsLayout ws_layout(int32_t R) {
    WsLayout L;
    size_t o = 0;
    // [0]: max h, [1]: min h; [64, 64 + kHmaxBlocks): per-block maxima, then
    // kHmaxBlocks per-block minima
    L.hmax = o;
    o = align256(o + sizeof(float) * (64 + 2 * kHmaxBlocks));
    const size_t n = (size_t)R * R;
    L.grad_n = o;
    o = align256(o + sizeof(float) * 3 * n);
    L.grad_z = o;
    o = align256(o + sizeof(float) * n);
    L.grad_g = o;
    o = align256(o + sizeof(float) * 2 * n);
    L.end = o;
    return L;
}

// ---------------------------------------------------------------------------
// Sobel normals (model/caustics.py:215-221): replicate padding, kernels
// [[-1,0,1],[-2,0,2],[-1,0,1]] (x) and its transpose (y), scaled by
// -1/(8 delta), z = 1, normalised.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sobel_raw(const float* __restrict__ h, int R, int r, int c, float& gx, float& gy) {
    const int r0 = max(r - 1, 0), r2 = min(r + 1, R - 1), c0 = max(c - 1, 0), c2 = min(c + 1, R - 1);
    const float* a = h + r0 * R;
    const float* b = h + r * R;
    const float* d = h + r2 * R;
    gx = (a[c2] - a[c0]) + 2.0f * (b[c2] - b[c0]) + (d[c2] - d[c0]);
    gy = (d[c0] - a[c0]) + 2.0f * (d[c] - a[c]) + (d[c2] - a[c2]);
}

__global__ __launch_bounds__(256) void normals_kernel(const float* __restrict__ h, int R, float s,
                                                      float* __restrict__ n) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= R * R) return;
    const int r = k / R, c = k - r * R;
    float gx, gy;
    sobel_raw(h, R, r, c, gx, gy);
    const V3 v = normalize(v3(gx * s, gy * s, 1.0f));
    n[3 * k] = v.x;
    n[3 * k + 1] = v.y;
    n[3 * k + 2] = v.z;
}

// max and min of h in two stages: per-block extremes over a grid of
// kHmaxBlocks workgroups (one workgroup alone walks 2^16 heights
// latency-bound), then hmax_kernel over those.  part: gridDim.x maxima, then
// gridDim.x minima.
__global__ __launch_bounds__(256) void hmax_part_kernel(const float* __restrict__ h, int64_t n,
                                                        float* __restrict__ part) {
    __shared__ float wpart[4], wmin[4];
    float m = -INFINITY, mn = INFINITY;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
        m = fmaxf(m, h[k]);
        mn = fminf(mn, h[k]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmaxf(m, __shfl_xor(m, o));
        mn = fminf(mn, __shfl_xor(mn, o));
    }
    if ((threadIdx.x & 63) == 0) {
        wpart[threadIdx.x >> 6] = m;
        wmin[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = fmaxf(fmaxf(wpart[0], wpart[1]), fmaxf(wpart[2], wpart[3]));
        part[gridDim.x + blockIdx.x] = fminf(fminf(wmin[0], wmin[1]), fminf(wmin[2], wmin[3]));
    }
}

// parts: n per-block maxima, then n minima.  out[0] = the maximum, out[1] =
// the minimum (and copy[0..1] too, when copy is not null)
__global__ __launch_bounds__(1024) void hmax_kernel(const float* __restrict__ parts, int64_t n,
                                                    float* __restrict__ out, float* __restrict__ copy) {
    __shared__ float part[16], pmin[16];
    float m = -INFINITY, mn = INFINITY;
    for (int64_t k = threadIdx.x; k < n; k += 1024) {
        m = fmaxf(m, parts[k]);
        mn = fminf(mn, parts[n + k]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmaxf(m, __shfl_xor(m, o));
        mn = fminf(mn, __shfl_xor(mn, o));
    }
    if ((threadIdx.x & 63) == 0) {
        part[threadIdx.x >> 6] = m;
        pmin[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = -INFINITY, rn = INFINITY;
        for (int w = 0; w < 16; ++w) {
            r = fmaxf(r, part[w]);
            rn = fminf(rn, pmin[w]);
        }
        out[0] = r;
        out[1] = rn;
        if (copy) {
            copy[0] = r;
            copy[1] = rn;
        }
    }
}

SceneArgs scene_args(const float* h, const float* normals, const hf_scene* sc, const float* hmax_dev) {
    SceneArgs S;
    S.h = h;
    S.normals = normals;
    S.R = sc->R;
    S.delta = (float)(2.0 / (double)(sc->R - 1));
    S.inv_delta = (float)((double)(sc->R - 1) / 2.0);
    S.sensor_z = sc->sensor_z;
    S.hmax = 0.0f;  // hmax and hmin: read from hmax_dev[0..1] on the device
    S.hmin = 0.0f;
    (void)hmax_dev;
    return S;
}

// ---------------------------------------------------------------------------
// trace contract
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void trace_kernel(SceneArgs S, const float* __restrict__ hmax,
                                                    const float* __restrict__ org, const float* __restrict__ dir,
                                                    int64_t N, float* __restrict__ t, float* __restrict__ uv,
                                                    int64_t* __restrict__ obj, int64_t* __restrict__ prim) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    S.hmax = hmax[0];
    S.hmin = hmax[1];
    const V3 o = v3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
    const V3 d = v3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    const Hit hit = trace(S, o, d);
    t[i] = hit.obj >= 0 ? hit.t : kMissT;
    uv[2 * i] = hit.obj >= 0 ? hit.b1 : -1.0f;
    uv[2 * i + 1] = hit.obj >= 0 ? hit.b2 : -1.0f;
    obj[i] = hit.obj;
    prim[i] = hit.obj >= 0 ? hit.prim : -1;
}

// ---------------------------------------------------------------------------
// the photon path
// ---------------------------------------------------------------------------
struct D3 {
    double x, y, z;
};

struct EmitArgs {
    int32_t n_side, channels, kind;
    float z0, pitch, flux;
    V3 light;
    D3 c00, c10, c01, c11;  // point light: unit directions to the plane's corners
    double diff_scale;
    uint64_t key;
    float ior[HF_MAX_CHANNELS];
};

struct PathState {
    V3 o, d, Du, Dv, Dt, Dp;
};

// E.ior[ch] by selects: a dynamic index into the kernel-argument array would
// copy it to scratch memory first
__device__ __forceinline__ float ior_of(const EmitArgs& E, int ch) {
    float v = E.ior[0];
#pragma unroll
    for (int k = 1; k < HF_MAX_CHANNELS; ++k) v = ch == k ? E.ior[k] : v;
    return v;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    // SplitMix64 (Steele, Lea and Flood, OOPSLA 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float u24(uint64_t bits) { return (float)(uint32_t)(bits >> 40) * (1.0f / 16777216.0f); }

// The point-light emitter works in double: Frisvad's basis loses
// 1 / (1 + n_z) of relative accuracy for directions near -z (rays from a light
// above the glass), so the direction must be exact well below float rounding
// before the basis is formed; results are rounded to float once.

// model/utils.py:84-88: (sin((1 - t) w) p0 + sin(t w) p1) / sin(w), w = acos(clamp(p0 . p1))
__device__ __forceinline__ D3 slerp_d(D3 p0, D3 p1, double t) {
    const double c = p0.x * p1.x + p0.y * p1.y + p0.z * p1.z;
    const double w = acos(fmin(fmax(c, -1.0), 1.0));
    const double sw = sin(w), a = sin((1.0 - t) * w) / sw, b = sin(t * w) / sw;
    return D3{a * p0.x + b * p1.x, a * p0.y + b * p1.y, a * p0.z + b * p1.z};
}

// Frisvad 2012 orthonormal basis of a unit vector (model/utils.py:33-53)
__device__ __forceinline__ void frisvad_d(D3 n, D3& b1, D3& b2) {
    if (n.z < -1.0 + 1e-5) {
        b1 = D3{0.0, -1.0, 0.0};
        b2 = D3{-1.0, 0.0, 0.0};
        return;
    }
    const double a = 1.0 / (1.0 + n.z), b = -n.x * n.y * a;
    b1 = D3{1.0 - a * n.x * n.x, b, -n.x};
    b2 = D3{b, 1.0 - a * n.y * n.y, -n.y};
}

__device__ __forceinline__ V3 to_v3(D3 v, double s) { return v3((float)(s * v.x), (float)(s * v.y), (float)(s * v.z)); }

__device__ __forceinline__ PathState emit(const EmitArgs& E, int64_t i, int& ch) {
    ch = (int)(i % E.channels);
    const int64_t cell = i / E.channels;
    PathState s;
    if (E.kind == 1) {
        // model/caustics.py:26-70: slerp between the corner directions with
        // per-sample coefficients (a, b), Du = Dv = 0, scaled orthonormal Dtheta, Dphi
        const uint64_t ctr = (uint64_t)cell << 2;
        const double a = u24(splitmix64(E.key ^ (ctr | 0))), b = u24(splitmix64(E.key ^ (ctr | 1)));
        const D3 lower = slerp_d(E.c00, E.c10, a), upper = slerp_d(E.c01, E.c11, a);
        const D3 d = slerp_d(lower, upper, b);
        D3 b1, b2;
        frisvad_d(d, b1, b2);
        s.d = to_v3(d, 1.0);
        s.o = E.light;
        s.Dt = to_v3(b1, E.diff_scale);
        s.Dp = to_v3(b2, E.diff_scale);
        s.Du = v3(0.f, 0.f, 0.f);
        s.Dv = v3(0.f, 0.f, 0.f);
        return s;
    }
    const int row = (int)(cell / E.n_side), col = (int)(cell - (int64_t)row * E.n_side);
    s.o = v3(-1.0f + (float)(2 * col + 1) / (float)E.n_side, -1.0f + (float)(2 * row + 1) / (float)E.n_side, E.z0);
    s.d = v3(0.f, 0.f, -1.f);
    s.Du = v3(2.0f * E.pitch, 0.f, 0.f);
    s.Dv = v3(0.f, 2.0f * E.pitch, 0.f);
    s.Dt = v3(0.f, 0.f, 0.f);
    s.Dp = v3(0.f, 0.f, 0.f);
    return s;
}

// model/photon_differential.py:70-78: Du' = Du + t Dth - beta d,
// beta = ng . (Du + t Dth) / (ng . d)
__device__ __forceinline__ V3 transfer1(V3 ng, V3 d, float t, V3 Du, V3 Dt) {
    const V3 X = Du + t * Dt;
    const float beta = dot(ng, X) / dot(ng, d);
    return X - beta * d;
}

// refract_reflect_differentials (model/caustics.py:152-212) with dN = 0:
// cosi = clamp(d . ns) unflipped, N = +-ns, eta = etai / etat.
struct Refr {
    V3 N;
    float cosi, eta, sq, coso, a;
    bool flip, tir, clamped;
};

// The two index ratios of a photon's path, eo / ei (entering the glass) and
// ei / eo (leaving it), divided once per photon instead of at every bounce.
struct Etas {
    float enter, exit;
};
__device__ __forceinline__ Etas etas_of(float eo, float ei) { return Etas{eo / ei, ei / eo}; }

__device__ __forceinline__ Refr refr_setup(V3 d, V3 ns, const Etas& et) {
    Refr R;
    const float raw = dot(d, ns);
    R.clamped = !(raw >= -1.0f && raw <= 1.0f);
    R.cosi = fminf(fmaxf(raw, -1.0f), 1.0f);
    R.flip = R.cosi > 0.0f;
    R.N = R.flip ? -ns : ns;
    R.eta = R.flip ? et.exit : et.enter;
    const float k = 1.0f - R.eta * R.eta * (1.0f - R.cosi * R.cosi);
    R.tir = !(k > 0.0f);
    R.sq = R.tir ? 0.0f : sqrtf(k);
    R.coso = R.eta * R.cosi + R.sq;
    R.a = R.tir ? 0.0f : R.eta + R.eta * R.eta * R.cosi / R.sq;
    return R;
}

__device__ __forceinline__ void refract_apply(const Refr& R, V3& d, V3& Dt, V3& Dp) {
    if (!R.tir) {
        const V3 dn = R.eta * d - R.coso * R.N;
        Dt = R.eta * Dt - (R.a * dot(Dt, R.N)) * R.N;
        Dp = R.eta * Dp - (R.a * dot(Dp, R.N)) * R.N;
        d = dn;
    } else {
        const V3 dn = d - (2.0f * R.cosi) * R.N;
        Dt = Dt - (2.0f * dot(Dt, R.N)) * R.N;
        Dp = Dp - (2.0f * dot(Dp, R.N)) * R.N;
        d = dn;
    }
}

struct Splat {
    float Ep, x, y, radius;
    V3 M0, M1;
    bool good;
};

// model/caustics.py:224-239
__device__ __forceinline__ Splat splat_inputs(V3 p, V3 Du, V3 Dv, V3 ns, float flux) {
    Splat s;
    const V3 tmp = cross(Dv, ns);
    const float den = dot(Du, tmp);
    s.good = den != 0.0f;
    const float f = 2.0f / den;
    s.M0 = f * tmp;
    s.M1 = f * cross(ns, Du);
    s.radius = 0.5f * fmaxf(norm(Du), norm(Dv));
    s.Ep = flux / norm(cross(Du, Dv)) * 4.0f;
    s.x = p.x;
    s.y = p.y;
    return s;
}

struct OutArgs {
    float *Ep, *xp, *Mp, *radius;
    int64_t* cp;
    int64_t N;
};

__device__ __forceinline__ void write_dropped(const OutArgs& O, int64_t i) {
    const int64_t N = O.N;
    O.Ep[i] = 0.0f;
    O.xp[i] = 0.0f;
    O.xp[N + i] = 0.0f;
    for (int k = 0; k < 6; ++k) O.Mp[k * N + i] = 0.0f;
    O.radius[i] = 0.0f;
    O.cp[i] = -1;
}

// Hit records for the adjoint: the first kHitSlots hits of every photon that
// reaches the sensor, [kHitSlots][N] float4 (t, b1, b2, obj << 30 | prim), and
// per photon the number recorded -- or why the photon was dropped: 0 = it
// missed the scene, kHitsDegenerate = it reached the sensor but the splat's
// den != 0 test masked it, kHitsCapped = still bouncing after kMaxBounces
// hits; kHitsRetrace = more bounces than slots (the adjoint traces that path
// again).  Dropped photons get no gradient.  hf_hits_stats_get counts these
// codes when asked, so the simulate itself keeps no counters.
constexpr int kHitSlots = 4;
constexpr uint8_t kHitsDegenerate = 253;
constexpr uint8_t kHitsCapped = 254;
constexpr uint8_t kHitsRetrace = 255;
static_assert(kMaxBounces < kHitsDegenerate, "hit counts must stay below the marker values");

enum { kReached = 0, kMissed = 1, kDegenerate = 2, kCapped = 3 };

__device__ __forceinline__ float4 pack_hit(const Hit& h) {
    return make_float4(h.t, h.b1, h.b2, __int_as_float((int)(((uint32_t)h.obj << 30) | (uint32_t)h.prim)));
}

__device__ __forceinline__ Hit unpack_hit(float4 v) {
    const uint32_t op = (uint32_t)__float_as_int(v.w);
    return Hit{v.x, v.y, v.z, (int)(op >> 30), (int)(op & 0x3fffffffu)};
}

// One photon's path (model/caustics.py:242-291): emit, then per hit trace,
// advance, transfer and either splat (ground) or refract, for at most
// kMaxBounces hits.  Dropped photons get cp = -1 and, with hit records, the
// code of the reason.
__device__ __forceinline__ void simulate_photon(SceneArgs S, const float* __restrict__ hmax, const EmitArgs& E,
                                               const OutArgs& O, float4* __restrict__ hits,
                                               uint8_t* __restrict__ nhit, int64_t i) {
    S.hmax = hmax[0];
    S.hmin = hmax[1];
    int ch;
    PathState s = emit(E, i, ch);
    const Etas et = etas_of(1.0f, ior_of(E, ch));  // outside: index 1
    int nb = 0;
    int outcome = kCapped;  // unless the loop ends earlier
    for (int bounce = 0; bounce < kMaxBounces; ++bounce) {
        const Hit hit = trace(S, s.o, s.d);
        if (hit.obj < 0) {  // miss: advance_path drops it
            outcome = kMissed;
            break;
        }
        if (hits && nb < kHitSlots) hits[(int64_t)nb * O.N + i] = pack_hit(hit);
        ++nb;
        const float t = hit.t;
        const V3 p = s.o + t * s.d;
        const TriVerts T = tri_verts(S, hit.obj, hit.prim);
        const V3 ng = geometric_normal(T);
        const V3 ns = normalize(shading_raw(T, hit.b1, hit.b2));
        const V3 Du = transfer1(ng, s.d, t, s.Du, s.Dt);
        const V3 Dv = transfer1(ng, s.d, t, s.Dv, s.Dp);
        if (hit.obj == 1) {
            const Splat sp = splat_inputs(p, Du, Dv, ns, E.flux);
            if (!sp.good) {  // the reference's den != 0 mask (caustics.py:226-230)
                outcome = kDegenerate;
                break;
            }
            const int64_t N = O.N;
            O.Ep[i] = sp.Ep;
            O.xp[i] = sp.x;
            O.xp[N + i] = sp.y;
            O.Mp[i] = sp.M0.x;
            O.Mp[N + i] = sp.M0.y;
            O.Mp[2 * N + i] = sp.M0.z;
            O.Mp[3 * N + i] = sp.M1.x;
            O.Mp[4 * N + i] = sp.M1.y;
            O.Mp[5 * N + i] = sp.M1.z;
            O.radius[i] = sp.radius;
            O.cp[i] = ch;
            if (nhit) nhit[i] = nb <= kHitSlots ? (uint8_t)nb : kHitsRetrace;
            return;
        }
        const Refr R = refr_setup(s.d, ns, et);
        refract_apply(R, s.d, s.Dt, s.Dp);
        s.Du = Du;
        s.Dv = Dv;
        s.o = p;
    }
    write_dropped(O, i);
    if (nhit) nhit[i] = outcome == kCapped ? kHitsCapped : outcome == kDegenerate ? kHitsDegenerate : (uint8_t)0;
}

// hits, nhit: null, or the hit records to write.
// 5 wavefronts per SIMD (<= 102 VGPRs; a few spill): ~1.5 % faster per
// Landweber step than the 4 that its natural 103 registers allow
__global__ __launch_bounds__(256, 5) void simulate_kernel(SceneArgs S, const float* __restrict__ hmax, EmitArgs E,
                                                       OutArgs O, float4* __restrict__ hits,
                                                       uint8_t* __restrict__ nhit) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < O.N) simulate_photon(S, hmax, E, O, hits, nhit, i);
}

// Path outcomes from the hit-count bytes: per-workgroup LDS counters, one
// device atomic per workgroup and outcome into stats[kStatWords] (cleared by
// stats_zero_kernel before).
__global__ __launch_bounds__(256) void stats_zero_kernel(uint32_t* __restrict__ stats) {
    if (threadIdx.x < kStatWords) stats[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void path_stats_kernel(const uint8_t* __restrict__ nhit, int64_t N,
                                                         uint32_t* __restrict__ stats) {
    __shared__ uint32_t c[kStatWords];
    if (threadIdx.x < kStatWords) c[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t n = 0, miss = 0, deg = 0, cap = 0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += stride) {
        const uint8_t v = nhit[i];
        ++n;
        miss += v == 0 ? 1u : 0u;
        deg += v == kHitsDegenerate ? 1u : 0u;
        cap += v == kHitsCapped ? 1u : 0u;
    }
    atomicAdd(&c[kReached], n - miss - deg - cap);
    atomicAdd(&c[kMissed], miss);
    atomicAdd(&c[kDegenerate], deg);
    atomicAdd(&c[kCapped], cap);
    __syncthreads();
    if (threadIdx.x < kStatWords && c[threadIdx.x]) atomicAdd(&stats[threadIdx.x], c[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
struct Tape {
    V3 d, Du, Dv, Dt, Dp;
    float t, b1, b2;
    int obj, prim;
};

// One bounce of the forward path after its hit: move to the hit point,
// transfer the differentials and refract -- the update of simulate_kernel,
// operation for operation, so a replay from recorded hits is bit-identical.
__device__ __forceinline__ void advance(const SceneArgs& S, PathState& s, const Hit& hit, const Etas& et) {
    const V3 p = s.o + hit.t * s.d;
    const TriVerts T = tri_verts(S, hit.obj, hit.prim);
    const V3 ng = geometric_normal(T);
    const V3 ns = normalize(shading_raw(T, hit.b1, hit.b2));
    const V3 Du = transfer1(ng, s.d, hit.t, s.Du, s.Dt);
    const V3 Dv = transfer1(ng, s.d, hit.t, s.Dv, s.Dp);
    const Refr R = refr_setup(s.d, ns, et);
    refract_apply(R, s.d, s.Dt, s.Dp);
    s.Du = Du;
    s.Dv = Dv;
    s.o = p;
}

__device__ __forceinline__ Tape tape_of(const PathState& s, const Hit& h) {
    return Tape{s.d, s.Du, s.Dv, s.Dt, s.Dp, h.t, h.b1, h.b2, h.obj, h.prim};
}

__device__ __forceinline__ void atomic_add3(float* p, V3 v) {
    unsafeAtomicAdd(p, v.x);
    unsafeAtomicAdd(p + 1, v.y);
    unsafeAtomicAdd(p + 2, v.z);
}

// adjoint of Du' = X - beta d, X = Du + t Dt, beta = ng . X / (ng . d):
// adds to gDu, gDt, gd and (optionally) gng
__device__ __forceinline__ void transfer1_adj(V3 ng, V3 d, float t, V3 Du, V3 Dt, V3 g, V3& gDu, V3& gDt, V3& gd,
                                              V3& gng) {
    const V3 X = Du + t * Dt;
    const float den = dot(ng, d);
    const float beta = dot(ng, X) / den;
    const float gd_ = dot(g, d);
    // d(Du') = dX - dbeta d - beta dd;  dbeta = (ng.dX + X.dng - beta (ng.dd + d.dng)) / den
    const V3 gX = g - (gd_ / den) * ng;
    gDu = gDu + gX;
    gDt = gDt + t * gX;
    gd = gd + (-beta) * g + (gd_ * beta / den) * ng;
    gng = gng + (-gd_ / den) * (X - beta * d);
}

// Per-vertex gradient accumulator of one workgroup: a PxP patch of top
// vertices in LDS (the rays of a 16x16 ray tile land on a few cells), global
// atomics for vertices outside the patch.  The patch sums in double: gfx950's
// LDS float add serialises a wavefront's lanes (~190 cycles per instruction
// per CU) while its double add does not (~9), and the sum rounds to float
// once, at the flush.
constexpr int kPatch = 8;

struct PatchAcc {
    double* n;  // [kPatch][kPatch][3] in LDS
    double* z;  // [kPatch][kPatch] in LDS
    int r0, c0, R;
    float* grad_n;
    float* grad_z;
    __device__ __forceinline__ bool inside(int r, int c, int& k) const {
        const int dr = r - r0, dc = c - c0;
        k = dr * kPatch + dc;
        return dr >= 0 && dr < kPatch && dc >= 0 && dc < kPatch;
    }
    __device__ __forceinline__ void add_n(int r, int c, V3 v) const {
        int k;
        if (inside(r, c, k)) {
            atomicAdd(n + 3 * k, (double)v.x);
            atomicAdd(n + 3 * k + 1, (double)v.y);
            atomicAdd(n + 3 * k + 2, (double)v.z);
        } else {
            atomic_add3(grad_n + 3 * (r * R + c), v);
        }
    }
    __device__ __forceinline__ void add_z(int r, int c, float v) const {
        int k;
        if (inside(r, c, k))
            atomicAdd(z + k, (double)v);
        else
            unsafeAtomicAdd(grad_z + r * R + c, v);
    }
};

__device__ __forceinline__ void adjoint_photon(const SceneArgs& S, const EmitArgs& E, int64_t i, int64_t N,
                                               const float* __restrict__ gEp, const float* __restrict__ gxp,
                                               const float* __restrict__ gMp, const PatchAcc& A,
                                               const float4* __restrict__ hits, const uint8_t* __restrict__ nhit) {
    // with the forward's hit records the path is replayed without tracing and
    // without a tape: the state at a bounce is advanced from the emitted one
    const int m = nhit ? (int)nhit[i] : (int)kHitsRetrace;
    // dropped in the forward (0, kHitsCapped): no gradient; any other count
    // above the recorded slots but the re-trace marker cannot occur
    if (m == 0 || (m > kHitSlots && m != kHitsRetrace)) return;
    const bool replay = m != kHitsRetrace;
    int ch;
    const PathState s0 = emit(E, i, ch);
    const Etas et = etas_of(1.0f, ior_of(E, ch));  // outside: index 1
    auto hit_at = [&](int b) { return unpack_hit(hits[(int64_t)b * N + i]); };
    Tape tape[kMaxBounces];  // traced paths only
    int nb = 0;
    PathState s = s0, s1 = s0;  // s1: the state after the first bounce (replay)
    if (replay) {
        nb = m;  // the last recorded hit is the ground
        for (int b = 0; b + 1 < m; ++b) {
            advance(S, s, hit_at(b), et);
            if (b == 0) s1 = s;
        }
    } else {
        bool reached = false;
        // forward, recording
        for (int bounce = 0; bounce < kMaxBounces; ++bounce) {
            const Hit hit = trace(S, s.o, s.d);
            if (hit.obj < 0) break;
            tape[nb++] = tape_of(s, hit);
            if (hit.obj == 1) {
                const V3 p = s.o + hit.t * s.d;
                const TriVerts T = tri_verts(S, hit.obj, hit.prim);
                const V3 ng = geometric_normal(T);
                const V3 ns = normalize(shading_raw(T, hit.b1, hit.b2));
                const V3 Du = transfer1(ng, s.d, hit.t, s.Du, s.Dt);
                const V3 Dv = transfer1(ng, s.d, hit.t, s.Dv, s.Dp);
                const Splat sp = splat_inputs(p, Du, Dv, ns, E.flux);
                reached = sp.good;
                break;
            }
            advance(S, s, hit, et);
        }
        if (!reached) return;
    }
    // the path at bounce k (the tape, or a replay of the recorded hits)
    auto bounce_at = [&](int k) -> Tape {
        if (!replay) return tape[k];
        if (k == nb - 1) return tape_of(s, hit_at(k));
        if (k == 0) return tape_of(s0, hit_at(0));
        PathState r = s1;
        for (int b = 1; b < k; ++b) advance(S, r, hit_at(b), et);
        return tape_of(r, hit_at(k));
    };
    // upstream gradients of the splat inputs
    const float gE = gEp[i];
    const float gx0 = gxp[i], gx1 = gxp[N + i];
    const V3 gM0 = v3(gMp[i], gMp[N + i], gMp[2 * N + i]);
    const V3 gM1 = v3(gMp[3 * N + i], gMp[4 * N + i], gMp[5 * N + i]);
    // ---- last bounce: ground plane (constant geometry)
    const Tape L = bounce_at(nb - 1);
    V3 gp = v3(gx0, gx1, 0.0f), gd, gDu, gDv, gDt, gDp;
    {
        const TriVerts T = tri_verts(S, L.obj, L.prim);
        const V3 ng = geometric_normal(T);
        const V3 ns = normalize(shading_raw(T, L.b1, L.b2));
        const V3 Du = transfer1(ng, L.d, L.t, L.Du, L.Dt);
        const V3 Dv = transfer1(ng, L.d, L.t, L.Dv, L.Dp);
        // splat adjoint (ns constant)
        const V3 tmp = cross(Dv, ns), sv = cross(ns, Du), c = cross(Du, Dv);
        const float den = dot(Du, tmp), f = 2.0f / den;
        const float Lc = norm(c), Ep = E.flux / Lc * 4.0f;
        const V3 gc = (-Ep / (Lc * Lc) * gE) * c;
        V3 gDu2 = cross(Dv, gc), gDv2 = cross(gc, Du);
        V3 gtmp = f * gM0;
        const V3 gs = f * gM1;
        const float gf = dot(gM0, tmp) + dot(gM1, sv);
        const float gden = -(2.0f / (den * den)) * gf;
        gDu2 = gDu2 + gden * tmp;
        gtmp = gtmp + gden * Du;
        gDv2 = gDv2 + cross(ns, gtmp);
        gDu2 = gDu2 + cross(gs, ns);
        // transfer adjoint at the ground (ng constant)
        gDu = v3(0.f, 0.f, 0.f), gDv = gDu, gDt = gDu, gDp = gDu, gd = gDu;
        V3 gng = gDu;
        transfer1_adj(ng, L.d, L.t, L.Du, L.Dt, gDu2, gDu, gDt, gd, gng);
        transfer1_adj(ng, L.d, L.t, L.Dv, L.Dp, gDv2, gDv, gDp, gd, gng);
        // p = o + t d
        gd = gd + L.t * gp;
    }
    // ---- earlier bounces, newest first: gradients arrive for (o' = p, d', Du', Dv', Dt', Dp')
    for (int k = nb - 2; k >= 0; --k) {
        const Tape B = bounce_at(k);
        const TriVerts T = tri_verts(S, B.obj, B.prim);
        const V3 graw = cross(T.v[1] - T.v[0], T.v[2] - T.v[0]);
        const float glen = norm(graw);
        const V3 ng = glen > 0.0f ? (1.0f / glen) * graw : v3(0.f, 0.f, 0.f);
        const float w0 = 1.0f - B.b1 - B.b2;
        const V3 m = shading_raw(T, B.b1, B.b2);
        const float mlen = norm(m);
        const V3 ns = mlen > 0.0f ? (1.0f / mlen) * m : v3(0.f, 0.f, 0.f);
        const Refr R = refr_setup(B.d, ns, et);
        // refraction adjoint: outputs d', Dt', Dp' (gradients gd, gDt, gDp)
        V3 gd_in = v3(0.f, 0.f, 0.f), gDt_in = gd_in, gDp_in = gd_in, gN = gd_in;
        float gcosi = 0.0f;
        if (!R.tir) {
            gd_in = R.eta * gd;
            const float gcoso = -dot(gd, R.N);
            gN = gN + (-R.coso) * gd;
            const float tN = dot(B.Dt, R.N), gtN = dot(gDt, R.N);
            const float pN = dot(B.Dp, R.N), gpN = dot(gDp, R.N);
            gDt_in = R.eta * gDt - (R.a * gtN) * R.N;
            gDp_in = R.eta * gDp - (R.a * gpN) * R.N;
            const float ga = -(tN * gtN) - (pN * gpN);
            gN = gN + (-R.a) * (gtN * B.Dt + tN * gDt) + (-R.a) * (gpN * B.Dp + pN * gDp);
            const float e2 = R.eta * R.eta;
            const float dadc = e2 * (1.0f - e2) / (R.sq * R.sq * R.sq);
            gcosi = R.a * gcoso + dadc * ga;
        } else {
            gd_in = gd;
            gcosi = -2.0f * dot(gd, R.N);
            gN = gN + (-2.0f * R.cosi) * gd;
            const float tN = dot(B.Dt, R.N), gtN = dot(gDt, R.N);
            const float pN = dot(B.Dp, R.N), gpN = dot(gDp, R.N);
            gDt_in = gDt - (2.0f * gtN) * R.N;
            gDp_in = gDp - (2.0f * gpN) * R.N;
            gN = gN + (-2.0f) * (gtN * B.Dt + tN * gDt) + (-2.0f) * (gpN * B.Dp + pN * gDp);
        }
        V3 gns = R.flip ? -gN : gN;
        if (!R.clamped) {
            gd_in = gd_in + gcosi * ns;
            gns = gns + gcosi * B.d;
        }
        // transfer adjoint (Du', Dv' <- Du, Dv, Dt, Dp, d, ng)
        V3 gDu_in = v3(0.f, 0.f, 0.f), gDv_in = gDu_in, gng = gDu_in;
        transfer1_adj(ng, B.d, B.t, B.Du, B.Dt, gDu, gDu_in, gDt_in, gd_in, gng);
        transfer1_adj(ng, B.d, B.t, B.Dv, B.Dp, gDv, gDv_in, gDp_in, gd_in, gng);
        // p = o + t d: gradient of o is gp (passed on), of d is t gp
        gd_in = gd_in + B.t * gp;
        // shading normal -> vertex normals (top-surface triangles only: the
        // other faces' vertex normals are constants)
        const bool top_face = B.prim < prim_base(S.R).bottom;
        if (top_face && mlen > 0.0f) {
            const V3 gm = (1.0f / mlen) * (gns - dot(gns, ns) * ns);
            const float w[3] = {w0, B.b1, B.b2};
            for (int v = 0; v < 3; ++v)
                if (T.vr[v] >= 0) A.add_n(T.vr[v], T.vc[v], w[v] * gm);
        }
        // geometric normal -> vertex heights (vertices whose height is h: the
        // top surface's and the walls' upper ones; none on the bottom face)
        const bool any_h = T.vr[0] >= 0 || T.vr[1] >= 0 || T.vr[2] >= 0;
        if (any_h && glen > 0.0f) {
            const V3 gg = (1.0f / glen) * (gng - dot(gng, ng) * ng);
            const V3 e1 = T.v[1] - T.v[0], e2 = T.v[2] - T.v[0];
            const V3 ge1 = cross(e2, gg), ge2 = cross(gg, e1);
            const float gz[3] = {-ge1.z - ge2.z, ge1.z, ge2.z};
            for (int v = 0; v < 3; ++v)
                if (T.vr[v] >= 0) A.add_z(T.vr[v], T.vc[v], gz[v]);
        }
        gd = gd_in;
        gDu = gDu_in;
        gDv = gDv_in;
        gDt = gDt_in;
        gDp = gDp_in;
    }
}

// One workgroup per 16x16 tile of the emitter's ray grid (one channel):
// gradients of the vertices the tile's rays hit gather in LDS first.
__global__ __launch_bounds__(256) void adjoint_kernel(SceneArgs S, const float* __restrict__ hmax, EmitArgs E,
                                                      int64_t N, int tiles_x, const float* __restrict__ gEp,
                                                      const float* __restrict__ gxp, const float* __restrict__ gMp,
                                                      float* __restrict__ grad_n, float* __restrict__ grad_z,
                                                      const float4* __restrict__ hits,
                                                      const uint8_t* __restrict__ nhit) {
    __shared__ double pn[kPatch * kPatch * 3];
    __shared__ double pz[kPatch * kPatch];
    S.hmax = hmax[0];
    S.hmin = hmax[1];
    const int ch = blockIdx.y;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int col = tx * 16 + (threadIdx.x & 15), row = ty * 16 + (threadIdx.x >> 4);
    for (int k = threadIdx.x; k < kPatch * kPatch * 3; k += 256) pn[k] = 0.0;
    for (int k = threadIdx.x; k < kPatch * kPatch; k += 256) pz[k] = 0.0;
    PatchAcc A;
    A.n = pn;
    A.z = pz;
    A.R = S.R;
    A.grad_n = grad_n;
    A.grad_z = grad_z;
    // patch origin: one vertex before the cell under the tile's first ray centre
    const float x0 = -1.0f + (float)(2 * tx * 16 + 1) / (float)E.n_side;
    const float y0 = -1.0f + (float)(2 * ty * 16 + 1) / (float)E.n_side;
    A.c0 = (int)floorf((x0 + 1.0f) * S.inv_delta) - 1;
    A.r0 = (int)floorf((y0 + 1.0f) * S.inv_delta) - 1;
    __syncthreads();
    if (col < E.n_side && row < E.n_side) {
        const int64_t i = ((int64_t)row * E.n_side + col) * E.channels + ch;
        adjoint_photon(S, E, i, N, gEp, gxp, gMp, A, hits, nhit);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kPatch * kPatch; k += 256) {
        const int r = A.r0 + k / kPatch, c = A.c0 + k % kPatch;
        if (r < 0 || r >= S.R || c < 0 || c >= S.R) continue;  // never accumulated: outside the grid
        const float vz = (float)pz[k];
        if (vz != 0.0f) unsafeAtomicAdd(grad_z + r * S.R + c, vz);
        const V3 vn = v3((float)pn[3 * k], (float)pn[3 * k + 1], (float)pn[3 * k + 2]);
        if (vn.x != 0.0f || vn.y != 0.0f || vn.z != 0.0f) atomic_add3(grad_n + 3 * (r * S.R + c), vn);
    }
}

__global__ __launch_bounds__(256) void zero_kernel(float* __restrict__ p, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) p[k] = 0.0f;
}

// per vertex: gradient of the raw Sobel sums (gx, gy) from the gradient of
// the normalised normal: n = v / |v|, v = (s gx, s gy, 1)
__global__ __launch_bounds__(256) void normals_adj_kernel(const float* __restrict__ h, int R, float s,
                                                          const float* __restrict__ grad_n,
                                                          float* __restrict__ grad_g) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= R * R) return;
    const int r = k / R, c = k - r * R;
    float gx, gy;
    sobel_raw(h, R, r, c, gx, gy);
    const V3 v = v3(gx * s, gy * s, 1.0f);
    const float len = norm(v);
    const V3 n = (1.0f / len) * v;
    const V3 g = v3(grad_n[3 * k], grad_n[3 * k + 1], grad_n[3 * k + 2]);
    const V3 gv = (1.0f / len) * (g - dot(g, n) * n);
    grad_g[2 * k] = s * gv.x;
    grad_g[2 * k + 1] = s * gv.y;
}

// dL/dh[r, c] = grad_z[r, c] + sum over the vertices whose (clamped) Sobel
// stencil reads h[r, c] of the stencil weight times their (ggx, ggy)
__global__ __launch_bounds__(256) void grad_h_kernel(int R, const float* __restrict__ grad_g,
                                                     const float* __restrict__ grad_z, float* __restrict__ grad_h) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= R * R) return;
    const int r = k / R, c = k - r * R;
    float acc = grad_z[k];
    const float w[3] = {1.0f, 2.0f, 1.0f};
    for (int vr = max(r - 1, 0); vr <= min(r + 1, R - 1); ++vr)
        for (int vc = max(c - 1, 0); vc <= min(c + 1, R - 1); ++vc) {
            const float ggx = grad_g[2 * (vr * R + vc)], ggy = grad_g[2 * (vr * R + vc) + 1];
            for (int a = 0; a < 3; ++a) {
                // x stencil: rows vr - 1 + a (clamped), columns vc +- 1 (clamped)
                const int rr = min(max(vr - 1 + a, 0), R - 1);
                if (rr == r) {
                    if (min(vc + 1, R - 1) == c) acc += w[a] * ggx;
                    if (max(vc - 1, 0) == c) acc -= w[a] * ggx;
                }
                // y stencil: columns vc - 1 + a (clamped), rows vr +- 1 (clamped)
                const int cc = min(max(vc - 1 + a, 0), R - 1);
                if (cc == c) {
                    if (min(vr + 1, R - 1) == r) acc += w[a] * ggy;
                    if (max(vr - 1, 0) == r) acc -= w[a] * ggy;
                }
            }
        }
    grad_h[k] = acc;
}

// ---------------------------------------------------------------------------
// Landweber data term and update
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 256;    // workgroups of the Landweber update's volume sum
constexpr int kResidBlocks = 1024;  // workgroups of the residual (enough loads in flight for 1024^2 pixels)

__global__ __launch_bounds__(256) void residual_kernel(const float* __restrict__ img, const float* __restrict__ tgt,
                                                       int64_t n, float* __restrict__ resid,
                                                       float* __restrict__ partial) {
    __shared__ float part[4];
    float acc = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
        const float r = img[k] - tgt[k];
        resid[k] = r;
        acc = fmaf(r, r, acc);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const float* __restrict__ partial, int n, float scale,
                                                           float* __restrict__ out) {
    __shared__ float part[4];
    float acc = 0.0f;
    for (int k = threadIdx.x; k < n; k += 256) acc += partial[k];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = scale * ((part[0] + part[1]) + (part[2] + part[3]));
}

struct LwArgs {
    float tau, alpha, lower, upper, d1, d2, gamma;
    int32_t radius;
};

// x = mask * extShrink(d - tau g, tau alpha, lower, upper), written to x;
// per-block partial sums of x
__global__ __launch_bounds__(256) void landweber_step_kernel(const float* __restrict__ d,
                                                             const float* __restrict__ g,
                                                             const float* __restrict__ mask, int64_t n, LwArgs A,
                                                             float* __restrict__ x, float* __restrict__ partial) {
    __shared__ float part[4];
    float acc = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
        const float y = d[k] - A.tau * g[k];
        const float sgn = y > 0.0f ? 1.0f : (y < 0.0f ? -1.0f : 0.0f);
        float v = fmaxf(fabsf(y) - A.tau * A.alpha, 0.0f) * sgn;
        v = fmaxf(fminf(v, A.upper), A.lower);
        v *= mask[k];
        x[k] = v;
        acc += v;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// VconsHeur: box mean of x over (2 radius + 1)^2 (zero padding, divisor the
// full window, as avg_pool2d counts padding), then x +- mask mean sin(..) gamma
__global__ __launch_bounds__(256) void vcons_kernel(const float* __restrict__ x, const float* __restrict__ mask,
                                                    int R, LwArgs A, const float* __restrict__ partial, int nparts,
                                                    float* __restrict__ d) {
    // the volume: every workgroup reduces the partials in the same fixed
    // order, so all of them see the same value
    __shared__ float part[4];
    float acc = 0.0f;
    for (int k = threadIdx.x; k < nparts; k += 256) acc += partial[k];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    const float vx = (part[0] + part[1]) + (part[2] + part[3]);
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= R * R) return;
    const int r = k / R, c = k - r * R;
    float f = 0.0f;
    if (vx <= A.d1)
        f = sinf(vx / A.d1 * 3.14159265358979f) * A.gamma;
    else if (vx >= A.d2)
        f = -sinf((vx / A.d2 - 1.0f) * 3.14159265358979f) * A.gamma;
    float out = x[k];
    if (f != 0.0f) {
        float s = 0.0f;
        for (int rr = r - A.radius; rr <= r + A.radius; ++rr)
            for (int cc = c - A.radius; cc <= c + A.radius; ++cc)
                if (rr >= 0 && rr < R && cc >= 0 && cc < R) s += x[rr * R + cc];
        const int w = 2 * A.radius + 1;
        out += mask[k] * (s / (float)(w * w)) * f;
    }
    d[k] = out;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// hmax[0] = max and hmax[1] = min of the n heights (hmax: the workspace's
// hmax region); also written to copy[0..1] when copy is not null
void launch_hmax(const float* h, int64_t n, float* hmax, hipStream_t s, float* copy = nullptr) {
    hipLaunchKernelGGL(hmax_part_kernel, dim3(kHmaxBlocks), dim3(256), 0, s, h, n, hmax + 64);
    hipLaunchKernelGGL(hmax_kernel, dim3(1), dim3(1024), 0, s, (const float*)(hmax + 64), (int64_t)kHmaxBlocks, hmax,
                       copy);
}

EmitArgs emit_args(const hf_scene* sc, const hf_emitter* em) {
    EmitArgs E;
    E.n_side = em->n_side;
    E.channels = sc->channels;
    E.kind = em->kind;
    E.z0 = em->z0;
    E.pitch = (float)(2.0 / (double)em->n_side);
    E.flux = em->kind == 1 ? em->flux_total
                           : (float)((double)em->flux_total / ((double)em->n_side * (double)em->n_side));
    const double L[3] = {em->light[0], em->light[1], em->light[2]};
    auto corner = [&](double x, double y) {
        const double dx = x - L[0], dy = y - L[1], dz = (double)em->plane_z - L[2];
        const double n = std::sqrt(dx * dx + dy * dy + dz * dz);
        return D3{dx / n, dy / n, dz / n};
    };
    E.light = V3{em->light[0], em->light[1], em->light[2]};
    E.c00 = corner(-1, -1);
    E.c10 = corner(1, -1);
    E.c01 = corner(-1, 1);
    E.c11 = corner(1, 1);
    E.diff_scale = em->diff_scale;
    E.key = (uint64_t)em->seed * 0x100000001B3ull;
    for (int k = 0; k < HF_MAX_CHANNELS; ++k) E.ior[k] = sc->ior[k];
    return E;
}

// Size limits: 3 R^2 normal components and N = n_side^2 channels photons
// stay below 2^31, so 32-bit vertex indices and one-dimensional launches of
// one thread per photon hold.
constexpr int32_t kMaxR = 16384;
constexpr int32_t kMaxSide = 32768;
constexpr int64_t kMaxPhotons = ((int64_t)1 << 31) - 1;

int check_scene(const hf_scene* sc, const void* h, const void* normals) {
    if (!sc) return fail(PDS_ERR_ARGUMENT, "null scene");
    if (sc->R < 2 || sc->R > kMaxR) return fail(PDS_ERR_ARGUMENT, "scene R must be in [2, 16384]");
    if (sc->channels < 1 || sc->channels > HF_MAX_CHANNELS) return fail(PDS_ERR_ARGUMENT, "channels must be in [1, 8]");
    if (!h || !normals) return fail(PDS_ERR_ARGUMENT, "null height field or normals");
    return PDS_OK;
}

int check_emitter(const hf_emitter* em) {
    if (!em) return fail(PDS_ERR_ARGUMENT, "null emitter");
    if (em->n_side < 1 || em->n_side > kMaxSide) return fail(PDS_ERR_ARGUMENT, "emitter n_side must be in [1, 32768]");
    if (em->kind != 0 && em->kind != 1) return fail(PDS_ERR_ARGUMENT, "emitter kind must be 0 (parallel) or 1 (point)");
    return PDS_OK;
}

int hip_status(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
        return PDS_ERR_HIP;
    }
    return PDS_OK;
}

// hit buffer: records [kHitSlots][N] float4, counts [N] bytes, the maximum
// and minimum height of the field the records were traced on (the adjoint's
// re-traced paths use them, so they are not computed again), and room for
// the path counters that hf_hits_stats_get counts
struct HitsLayout {
    size_t recs, count, hmax, stats, end;
};

HitsLayout hits_layout(int64_t N) {
    HitsLayout L;
    L.recs = 0;
    L.count = (sizeof(float4) * kHitSlots * (size_t)N + 255) & ~(size_t)255;
    L.hmax = (L.count + (size_t)N + 255) & ~(size_t)255;
    L.stats = L.hmax + 16;  // after the max and min height
    L.end = L.stats + sizeof(uint32_t) * kStatWords;
    return L;
}

int simulate_impl(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em, float* Ep,
                  float* xp, float* Mp, int64_t* cp, float* radius, void* ws, size_t ws_bytes, void* hits,
                  size_t hits_bytes, bool want_hits, void* stream) {
    g_err[0] = '\0';
    int rc = check_scene(sc, h, normals);
    if (rc) return rc;
    if ((rc = check_emitter(em))) return rc;
    if (!Ep || !xp || !Mp || !cp || !radius) return fail(PDS_ERR_ARGUMENT, "null output buffer");
    const WsLayout L = ws_layout(sc->R);
    if (!ws || ws_bytes < L.end) return fail(PDS_ERR_WORKSPACE, "workspace smaller than hf_workspace_bytes(R)");
    const int64_t N = (int64_t)em->n_side * em->n_side * sc->channels;
    if (N > kMaxPhotons) return fail(PDS_ERR_CAPACITY, "n_side^2 * channels photons exceed 2^31 - 1");
    const HitsLayout HL = hits_layout(N);
    if (want_hits && (!hits || hits_bytes < HL.end))
        return fail(PDS_ERR_WORKSPACE, "hit buffer smaller than hf_hits_bytes");
    float* hmax = (float*)((char*)ws + L.hmax);
    OutArgs O{Ep, xp, Mp, radius, cp, N};
    float4* hrec = want_hits ? (float4*)((char*)hits + HL.recs) : nullptr;
    uint8_t* hcnt = want_hits ? (uint8_t*)hits + HL.count : nullptr;
    hipStream_t s = (hipStream_t)stream;
    launch_hmax(h, (int64_t)sc->R * sc->R, hmax, s, want_hits ? (float*)((char*)hits + HL.hmax) : nullptr);
    hipLaunchKernelGGL(simulate_kernel, dim3(blocks_for(N)), dim3(256), 0, s, scene_args(h, normals, sc, hmax),
                       (const float*)hmax, emit_args(sc, em), O, hrec, hcnt);
    return hip_status("hf_simulate");
}

int adjoint_impl(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em,
                 const float* grad_Ep, const float* grad_xp, const float* grad_Mp, float* grad_h, void* ws,
                 size_t ws_bytes, const void* hits, size_t hits_bytes, bool want_hits, void* stream) {
    g_err[0] = '\0';
    int rc = check_scene(sc, h, normals);
    if (rc) return rc;
    if ((rc = check_emitter(em))) return rc;
    if (!grad_Ep || !grad_xp || !grad_Mp || !grad_h) return fail(PDS_ERR_ARGUMENT, "null gradient buffer");
    const WsLayout L = ws_layout(sc->R);
    if (!ws || ws_bytes < L.end) return fail(PDS_ERR_WORKSPACE, "workspace smaller than hf_workspace_bytes(R)");
    const int64_t N = (int64_t)em->n_side * em->n_side * sc->channels;
    if (N > kMaxPhotons) return fail(PDS_ERR_CAPACITY, "n_side^2 * channels photons exceed 2^31 - 1");
    const HitsLayout HL = hits_layout(N);
    if (want_hits && (!hits || hits_bytes < HL.end))
        return fail(PDS_ERR_WORKSPACE, "hit buffer smaller than hf_hits_bytes");
    char* b = (char*)ws;
    float* hmax = (float*)(b + L.hmax);
    float* gn = (float*)(b + L.grad_n);
    float* gz = (float*)(b + L.grad_z);
    float* gg = (float*)(b + L.grad_g);
    const int R = sc->R;
    const int64_t RR = (int64_t)R * R;
    const float s_sobel = (float)(-1.0 / (8.0 * (2.0 / (double)(R - 1))));
    const float4* hrec = want_hits ? (const float4*)((const char*)hits + HL.recs) : nullptr;
    const uint8_t* hcnt = want_hits ? (const uint8_t*)hits + HL.count : nullptr;
    hipStream_t s = (hipStream_t)stream;
    // grad_n and grad_z are adjacent in the workspace: one fill clears both (and the padding between)
    const int64_t nz = (int64_t)((L.grad_z - L.grad_n) / sizeof(float)) + RR;
    hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(nz)), dim3(256), 0, s, gn, nz);
    // with hit records, the maximum height the simulate stored beside them
    const float* hm = hmax;
    if (want_hits)
        hm = (const float*)((const char*)hits + HL.hmax);
    else
        launch_hmax(h, RR, hmax, s);
    const int tiles_x = (em->n_side + 15) / 16;
    hipLaunchKernelGGL(adjoint_kernel, dim3((unsigned)(tiles_x * tiles_x), (unsigned)sc->channels), dim3(256), 0, s,
                       scene_args(h, normals, sc, hm), hm, emit_args(sc, em), N, tiles_x, grad_Ep, grad_xp, grad_Mp, gn,
                       gz, hrec, hcnt);
    hipLaunchKernelGGL(normals_adj_kernel, dim3(blocks_for(RR)), dim3(256), 0, s, h, R, s_sobel, (const float*)gn, gg);
    hipLaunchKernelGGL(grad_h_kernel, dim3(blocks_for(RR)), dim3(256), 0, s, R, (const float*)gg, (const float*)gz,
                       grad_h);
    return hip_status("hf_adjoint");
}

}  // namespace

}  // namespace hf

using namespace hf;

extern "C" {

const char* hf_last_error(void) { return g_err; }

size_t hf_workspace_bytes(int32_t R) { return R < 2 || R > kMaxR ? 0 : ws_layout(R).end; }

int hf_normals_f32(const float* h, int32_t R, float* normals, void* stream) {
    g_err[0] = '\0';
    if (R < 2 || R > kMaxR || !h || !normals) return fail(PDS_ERR_ARGUMENT, "invalid height field (R must be in [2, 16384])");
    const float s = (float)(-1.0 / (8.0 * (2.0 / (double)(R - 1))));
    hipLaunchKernelGGL(normals_kernel, dim3(blocks_for((int64_t)R * R)), dim3(256), 0, (hipStream_t)stream, h, R, s,
                       normals);
    return hip_status("hf_normals");
}

int hf_trace_f32(const float* h, const float* normals, const hf_scene* sc, const float* origins, const float* dirs,
                 int64_t N, float* t, float* uv, int64_t* object, int64_t* primitive, void* ws, size_t ws_bytes,
                 void* stream) {
    g_err[0] = '\0';
    int rc = check_scene(sc, h, normals);
    if (rc) return rc;
    if (N < 0) return fail(PDS_ERR_ARGUMENT, "negative ray count");
    if (N == 0) return PDS_OK;
    if (!origins || !dirs || !t || !uv || !object || !primitive) return fail(PDS_ERR_ARGUMENT, "null ray buffer");
    const WsLayout L = ws_layout(sc->R);
    if (!ws || ws_bytes < L.end) return fail(PDS_ERR_WORKSPACE, "workspace smaller than hf_workspace_bytes(R)");
    float* hmax = (float*)((char*)ws + L.hmax);
    hipStream_t s = (hipStream_t)stream;
    launch_hmax(h, (int64_t)sc->R * sc->R, hmax, s);
    hipLaunchKernelGGL(trace_kernel, dim3(blocks_for(N)), dim3(256), 0, s, scene_args(h, normals, sc, hmax),
                       (const float*)hmax, origins, dirs, N, t, uv, object, primitive);
    return hip_status("hf_trace");
}

size_t hf_hits_bytes(const hf_scene* sc, const hf_emitter* em) {
    if (!sc || !em || sc->channels < 1 || sc->channels > HF_MAX_CHANNELS || em->n_side < 1 || em->n_side > kMaxSide)
        return 0;
    const int64_t N = (int64_t)em->n_side * em->n_side * sc->channels;
    return N > kMaxPhotons ? 0 : hits_layout(N).end;
}

int hf_simulate_f32(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em, float* Ep,
                    float* xp, float* Mp, int64_t* cp, float* radius, void* ws, size_t ws_bytes, void* stream) {
    return simulate_impl(h, normals, sc, em, Ep, xp, Mp, cp, radius, ws, ws_bytes, nullptr, 0, false, stream);
}

int hf_simulate_hits_f32(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em, float* Ep,
                         float* xp, float* Mp, int64_t* cp, float* radius, void* ws, size_t ws_bytes, void* hits,
                         size_t hits_bytes, void* stream) {
    return simulate_impl(h, normals, sc, em, Ep, xp, Mp, cp, radius, ws, ws_bytes, hits, hits_bytes, true, stream);
}

int hf_hits_stats_get(void* hits, size_t hits_bytes, const hf_scene* sc, const hf_emitter* em, hf_path_stats* out,
                      void* stream) {
    g_err[0] = '\0';
    if (!out) return fail(PDS_ERR_ARGUMENT, "null stats pointer");
    const size_t need = hf_hits_bytes(sc, em);
    if (need == 0) return fail(PDS_ERR_ARGUMENT, "invalid scene or emitter");
    if (!hits || hits_bytes < need) return fail(PDS_ERR_WORKSPACE, "hit buffer smaller than hf_hits_bytes");
    const int64_t N = (int64_t)em->n_side * em->n_side * sc->channels;
    const HitsLayout HL = hits_layout(N);
    uint32_t* st = (uint32_t*)((char*)hits + HL.stats);
    uint32_t c[kStatWords];
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(stats_zero_kernel, dim3(1), dim3(256), 0, s, st);
    hipLaunchKernelGGL(path_stats_kernel, dim3(256), dim3(256), 0, s, (const uint8_t*)hits + HL.count, N, st);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(c, st, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return hip_status("hf_hits_stats_get");
    out->reached = c[kReached];
    out->missed = c[kMissed];
    out->degenerate = c[kDegenerate];
    out->capped = c[kCapped];
    for (int k = 0; k < 4; ++k) out->reserved[k] = 0;
    return PDS_OK;
}

int hf_adjoint_f32(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em,
                   const float* grad_Ep, const float* grad_xp, const float* grad_Mp, float* grad_h, void* ws,
                   size_t ws_bytes, void* stream) {
    return adjoint_impl(h, normals, sc, em, grad_Ep, grad_xp, grad_Mp, grad_h, ws, ws_bytes, nullptr, 0, false,
                        stream);
}

int hf_adjoint_hits_f32(const float* h, const float* normals, const hf_scene* sc, const hf_emitter* em,
                        const float* grad_Ep, const float* grad_xp, const float* grad_Mp, float* grad_h, void* ws,
                        size_t ws_bytes, const void* hits, size_t hits_bytes, void* stream) {
    return adjoint_impl(h, normals, sc, em, grad_Ep, grad_xp, grad_Mp, grad_h, ws, ws_bytes, hits, hits_bytes, true,
                        stream);
}

size_t hf_residual_workspace_bytes(int64_t n) { return n < 0 ? 0 : sizeof(float) * kResidBlocks; }

int hf_residual_f32(const float* img, const float* target, int64_t n, float* resid, float* loss, void* ws,
                    size_t ws_bytes, void* stream) {
    g_err[0] = '\0';
    if (n < 0 || (n > 0 && (!img || !target || !resid)) || !loss) return fail(PDS_ERR_ARGUMENT, "invalid residual args");
    if (!ws || ws_bytes < sizeof(float) * kResidBlocks) return fail(PDS_ERR_WORKSPACE, "residual workspace too small");
    hipStream_t s = (hipStream_t)stream;
    float* partial = (float*)ws;
    hipLaunchKernelGGL(residual_kernel, dim3(kResidBlocks), dim3(256), 0, s, img, target, n, resid, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, (const float*)partial, kResidBlocks, 0.5f, loss);
    return hip_status("hf_residual");
}

size_t hf_landweber_workspace_bytes(int32_t R) {
    return R < 1 || R > kMaxR ? 0 : align256(sizeof(float) * kRedBlocks) + sizeof(float) * (size_t)R * R;
}

int hf_landweber_f32(float* d, const float* grad, const float* mask, int32_t R, const hf_landweber* p, void* ws,
                     size_t ws_bytes, void* stream) {
    g_err[0] = '\0';
    if (R < 1 || R > kMaxR || !d || !grad || !mask || !p) return fail(PDS_ERR_ARGUMENT, "invalid Landweber args");
    if (p->volume_radius < 0 || p->volume_radius > 64) return fail(PDS_ERR_ARGUMENT, "volume_radius must be in [0, 64]");
    if (!ws || ws_bytes < hf_landweber_workspace_bytes(R)) return fail(PDS_ERR_WORKSPACE, "Landweber workspace too small");
    hipStream_t s = (hipStream_t)stream;
    float* partial = (float*)ws;
    float* x = (float*)((char*)ws + align256(sizeof(float) * kRedBlocks));
    LwArgs A{p->tau, p->alpha, p->lower, p->upper, p->d1, p->d2, p->gamma, p->volume_radius};
    const int64_t n = (int64_t)R * R;
    hipLaunchKernelGGL(landweber_step_kernel, dim3(kRedBlocks), dim3(256), 0, s, (const float*)d, grad, mask, n, A, x,
                       partial);
    hipLaunchKernelGGL(vcons_kernel, dim3(blocks_for(n)), dim3(256), 0, s, (const float*)x, mask, R, A,
                       (const float*)partial, kRedBlocks, d);
    return hip_status("hf_landweber");
}

}  // extern "C"
