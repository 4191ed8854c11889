// hf_device.hpp -- device geometry of the height-field caustic pipeline:
// the substrate scene, a grid-traversal ray tracer with the reference's
// trace contract, shading, differential transfer and refraction.
//
// Scene (model/renderable_object.py:81-193, :249-308): object 0 is the glass
// substrate over [-1, 1]^2 -- its top is the height field h (R x R vertices,
// two triangles per cell), its bottom the plane z = 0, and four side walls
// join them; object 1 is the ground plane z = sensor_z over [-1, 1]^2 (two
// triangles).  Triangle order inside object 0: top 2(R-1)^2, bottom
// 2(R-1)^2, front (y = -1), back (y = +1), left (x = -1), right (x = +1),
// 2(R-1) each -- the order update_from_height_field builds, so primitive
// indices match the reference's.
//
// Trace contract (PyOptix/RayTrace.cpp:233-291, kernel/ray_programs.cu:84-101):
// closest hit with t >= 1e-3 in units of the (possibly non-unit) direction,
// barycentrics (b1, b2) with hit = (1 - b1 - b2) v0 + b1 v1 + b2 v2, object
// and primitive index; a miss gives t = -2, b = (-1, -1), indices -1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hf.h"

namespace hf {

constexpr float kTMin = 1e-3f;
constexpr float kCullZ = 1e-3f;  // relative height margin of the cell cull (float rounding of z: ~1e-7 relative)
constexpr float kMissT = -2.0f;
// Hits a photon path may take (HF_MAX_BOUNCES of hf.h).  The reference loops
// until every photon has left the scene (model/caustics.py:255); a GPU thread
// needs a bound, and the FP64 oracle uses the same one: a photon still
// bouncing after this many hits is dropped (and counted).
constexpr int kMaxBounces = HF_MAX_BOUNCES;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float norm(V3 a) { return sqrtf(dot(a, a)); }
// model/utils.py:25-31: v / |v| for |v| > 0, else 0
__device__ __forceinline__ V3 normalize(V3 a) {
    const float n = norm(a);
    return n > 0.0f ? (1.0f / n) * a : v3(0.f, 0.f, 0.f);
}

struct SceneArgs {
    const float* h;        // [R, R] heights, row-major (row = y, column = x)
    const float* normals;  // [R, R, 3] unit Sobel normals of the top surface
    int32_t R;
    float delta;           // 2 / (R - 1)
    float inv_delta;
    float sensor_z;        // ground plane height
    float hmax;            // max(h), min(h): the top surface's z range (traversal bounds)
    float hmin;
};

__device__ __forceinline__ float height(const SceneArgs& S, int r, int c) { return S.h[r * S.R + c]; }
__device__ __forceinline__ float coord(const SceneArgs& S, int k) { return -1.0f + (float)k * S.delta; }
__device__ __forceinline__ V3 top_pt(const SceneArgs& S, int r, int c) {
    return v3(coord(S, c), coord(S, r), height(S, r, c));
}
__device__ __forceinline__ V3 bot_pt(const SceneArgs& S, int r, int c) { return v3(coord(S, c), coord(S, r), 0.0f); }

// Primitive index bases inside object 0.
struct PrimBase {
    int top, bottom, front, back, left, right;
};
__device__ __forceinline__ PrimBase prim_base(int R) {
    const int cells = (R - 1) * (R - 1), e = R - 1;
    PrimBase b;
    b.top = 0;
    b.bottom = 2 * cells;
    b.front = 4 * cells;
    b.back = b.front + 2 * e;
    b.left = b.back + 2 * e;
    b.right = b.left + 2 * e;
    return b;
}

// Where a vertex of a triangle comes from: a top-surface grid vertex (its
// height and normal depend on h) or a constant.
struct TriVerts {
    V3 v[3];
    V3 n[3];        // vertex normals
    int vr[3], vc[3];  // grid vertex of a top vertex, vr = -1 for constant vertices
};

constexpr int tri_code(int a0, int a1, int a2, int b0, int b1, int b2) {
    return a0 | a1 << 2 | a2 << 4 | b0 << 6 | b1 << 8 | b2 << 10;
}
// Every primitive of object 0 is one of the two triangles of a quad.  Corner
// order of a grid quad: q0 = (r, c), q1 = (r, c + 1), q2 = (r + 1, c),
// q3 = (r + 1, c + 1); triangles (model/renderable_object.py:52-78, the two
// halves of prim / 2 in grid_shuffle order):
//   top    (r,c) (r,c+1) (r+1,c)      and (r,c+1) (r+1,c+1) (r+1,c)
//   bottom (r,c) (r+1,c) (r,c+1)      and (r,c+1) (r+1,c) (r+1,c+1)
// A code holds per triangle three 2-bit corner indices (triangle 0 in bits
// 0-5, triangle 1 in bits 6-11, vertex k in bits 2k, 2k+1).
constexpr int kTopCode = tri_code(0, 1, 2, 1, 3, 2);
constexpr int kBottomCode = tri_code(0, 2, 1, 1, 2, 3);
// Corner order of a wall segment between grid vertices A and B (B one step
// further along the wall): q0 = A at z = 0, q1 = A on the top surface,
// q2 = B at z = 0, q3 = B on the top surface.  Per wall, in the order
// left x = -1, right x = +1, front y = -1, back y = +1:
constexpr int kWallCode[4] = {tri_code(0, 1, 2, 1, 3, 2), tri_code(0, 2, 1, 1, 2, 3), tri_code(0, 2, 1, 2, 3, 1),
                              tri_code(0, 1, 2, 2, 1, 3)};

__device__ __forceinline__ int pick4i(int a, int b, int c, int d, int k) {
    const int lo = (k & 1) ? b : a, hi = (k & 1) ? d : c;
    return (k & 2) ? hi : lo;
}

// Vertices and vertex normals of (obj, prim) -- grid_shuffle order
// (model/renderable_object.py:52-78).  The primitive is described by its
// quad (four corner grid vertices, which of them lie on the top surface, the
// face normal unless the top surface's vertex normals apply) and its corner
// code; the three vertices are then picked with constant array indices, so
// the result stays in registers.
__device__ __forceinline__ TriVerts tri_verts(const SceneArgs& S, int obj, int prim) {
    TriVerts T;
    if (obj == 1) {  // ground plane
        const float z = S.sensor_z;
        if (prim == 0) {
            T.v[0] = v3(-1.f, -1.f, z), T.v[1] = v3(1.f, -1.f, z), T.v[2] = v3(-1.f, 1.f, z);
        } else {
            T.v[0] = v3(-1.f, 1.f, z), T.v[1] = v3(1.f, -1.f, z), T.v[2] = v3(1.f, 1.f, z);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T.n[k] = v3(0.f, 0.f, 1.f);
            T.vr[k] = -1;
            T.vc[k] = 0;
        }
        return T;
    }
    const int R = S.R, e = R - 1;
    const PrimBase B = prim_base(R);
    int r0, c0, r1, c1, r2, c2, r3, c3;  // corner grid vertices
    int topmask, code;                   // corners on the top surface (bit k: corner k)
    bool grid_normals = false;           // top surface: the vertices' own normals
    V3 fn = v3(0.f, 0.f, 0.f);           // else the face normal
    if (prim < B.front) {  // top or bottom: a grid cell
        const bool top = prim < B.bottom;
        const int cell = (top ? prim : prim - B.bottom) >> 1, r = cell / e, c = cell - r * e;
        r0 = r, c0 = c, r1 = r, c1 = c + 1, r2 = r + 1, c2 = c, r3 = r + 1, c3 = c + 1;
        topmask = top ? 15 : 0;
        code = top ? kTopCode : kBottomCode;
        grid_normals = top;
        if (!top) fn = v3(0.f, 0.f, -1.f);
    } else {  // a wall segment: corners A, A, B, B
        int w, k;
        if (prim < B.back) {
            w = 2, k = (prim - B.front) >> 1;
        } else if (prim < B.left) {
            w = 3, k = (prim - B.back) >> 1;
        } else if (prim < B.right) {
            w = 0, k = (prim - B.left) >> 1;
        } else {
            w = 1, k = (prim - B.right) >> 1;
        }
        const bool xw = w < 2;
        const int fixed = (w & 1) ? e : 0;  // column of an x wall, row of a y wall
        r0 = r1 = xw ? k : fixed;
        c0 = c1 = xw ? fixed : k;
        r2 = r3 = xw ? k + 1 : fixed;
        c2 = c3 = xw ? fixed : k + 1;
        topmask = 2 | 8;
        code = w == 0 ? kWallCode[0] : (w == 1 ? kWallCode[1] : (w == 2 ? kWallCode[2] : kWallCode[3]));
        const float sgn = (w & 1) ? 1.0f : -1.0f;
        fn = xw ? v3(sgn, 0.f, 0.f) : v3(0.f, sgn, 0.f);
    }
    const int sh = 6 * (prim & 1);  // quad halves are even/odd primitives: every base is even
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int q = (code >> (sh + 2 * k)) & 3;
        const int r = pick4i(r0, r1, r2, r3, q), c = pick4i(c0, c1, c2, c3, q);
        const bool on_top = (topmask >> q) & 1;
        T.v[k] = v3(coord(S, c), coord(S, r), on_top ? height(S, r, c) : 0.0f);
        if (grid_normals) {
            const float* nn = S.normals + 3 * (r * R + c);
            T.n[k] = v3(nn[0], nn[1], nn[2]);
        } else {
            T.n[k] = fn;
        }
        T.vr[k] = on_top ? r : -1;
        T.vc[k] = on_top ? c : 0;
    }
    return T;
}

struct Hit {
    float t, b1, b2;
    int obj, prim;
};

__device__ __forceinline__ Hit miss_hit() { return Hit{kMissT, -1.0f, -1.0f, -1, -1}; }

// Moeller-Trumbore; keeps the closer hit with t >= kTMin (no culling).
__device__ __forceinline__ void mt_test(V3 o, V3 d, V3 v0, V3 v1, V3 v2, int obj, int prim, Hit& best) {
    const V3 e1 = v1 - v0, e2 = v2 - v0;
    const V3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (det == 0.0f) return;
    const float inv = 1.0f / det;
    const V3 tv = o - v0;
    const float u = dot(tv, pv) * inv;
    if (!(u >= 0.0f)) return;
    const V3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (!(v >= 0.0f) || u + v > 1.0f) return;
    const float t = dot(e2, qv) * inv;
    if (!(t >= kTMin)) return;
    if (best.obj < 0 || t < best.t) best = Hit{t, u, v, obj, prim};
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ V3 pick4(const V3& q0, const V3& q1, const V3& q2, const V3& q3, int k) {
    const V3 lo = (k & 1) ? q1 : q0, hi = (k & 1) ? q3 : q2;
    return (k & 2) ? hi : lo;
}

// The two triangles of a quad of object 0, primitives prim0 and prim0 + 1:
// corners q0..q3, and per triangle three 2-bit corner indices (triangle 0 in
// bits 0-5, triangle 1 in bits 6-11, vertex k of a triangle in bits 2k, 2k+1).
// A compile-time code folds the corner selection away.
__device__ __forceinline__ void quad_test(V3 o, V3 d, const V3& q0, const V3& q1, const V3& q2, const V3& q3,
                                          int code, int prim0, Hit& best) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int k = code >> (6 * half);
        mt_test(o, d, pick4(q0, q1, q2, q3, k & 3), pick4(q0, q1, q2, q3, (k >> 2) & 3),
                pick4(q0, q1, q2, q3, (k >> 4) & 3), 0, prim0 + half, best);
    }
}

__device__ __forceinline__ bool same_hit(const Hit& a, const Hit& b) {
    return a.prim == b.prim && a.obj == b.obj && a.t == b.t;
}

// Closest hit of the ray (o, d) with the scene: among all triangles, the one
// with the smallest t >= kTMin, ties going to the first in the order ground,
// bottom, walls, top surface (the result of testing them one after another,
// a hit replacing the best only when strictly closer).  Hit values come from
// mt_test alone; the plane parameters, clipping and cell walk only choose
// which triangles to test, so they multiply by one hardware reciprocal per
// axis instead of dividing.
//
// The top surface is walked first.  A plane (ground, bottom, a wall) whose
// intersection parameter lies clearly beyond a hit already found -- by a
// relative 1e-5, far above the ~1e-7 rounding of t -- cannot hold the
// closest hit or tie with it, so its triangles are not tested; the result is
// the same as testing everything.  (The bottom's and the walls' fallback to
// neighbouring triangles compares coplanar triangles, whose t agree to
// rounding, so it too is unaffected.)
__device__ inline Hit trace(const SceneArgs& S, V3 o, V3 d) {
    const int R = S.R, e = R - 1;
    const PrimBase B = prim_base(R);
    // hardware reciprocals (1 ulp): they only steer the walk and the skips,
    // whose margins are far wider; inf on a zero component, never used then
    const float rdx = __builtin_amdgcn_rcpf(d.x), rdy = __builtin_amdgcn_rcpf(d.y), rdz = __builtin_amdgcn_rcpf(d.z);
    Hit top = miss_hit();
    // top surface: walk the cells under the ray's xy path while its z lies in
    // the surface's range [hmin, hmax] (widened below); the first cell with a
    // hit holds the closest top hit.  Cells are visited in increasing t.
    {
        float t0 = 0.0f, t1 = INFINITY;
        // clip to the square [-1, 1]^2
        auto clip = [&](float oo, float dd, float rd) {
            if (dd == 0.0f) {
                if (oo < -1.0f || oo > 1.0f) t1 = -1.0f;
                return;
            }
            float a = (-1.0f - oo) * rd, b = (1.0f - oo) * rd;
            if (a > b) {
                const float tmp = a;
                a = b;
                b = tmp;
            }
            t0 = fmaxf(t0, a);
            t1 = fminf(t1, b);
        };
        clip(o.x, d.x, rdx);
        clip(o.y, d.y, rdy);
        // clip to the slab zlo <= z <= hmax: the top surface lies within
        // [hmin, hmax] (zlo = hmin less 1e-3 relative, far above rounding;
        // never below the bottom face's 0)
        const float zlo = fmaxf(0.0f, S.hmin - kCullZ * (1.0f + fabsf(S.hmin)));
        if (d.z == 0.0f) {
            if (o.z < zlo || o.z > S.hmax) t1 = -1.0f;
        } else {
            float a = (zlo - o.z) * rdz, b = (S.hmax - o.z) * rdz;
            if (a > b) {
                const float tmp = a;
                a = b;
                b = tmp;
            }
            t0 = fmaxf(t0, a);
            t1 = fminf(t1, b);
        }
        if (t0 <= t1 && t1 >= 0.0f) {
            // start cell (slightly inside the entry point) and DDA steps
            const float tm = t0;
            const float x = fmaf(tm, d.x, o.x), y = fmaf(tm, d.y, o.y);
            int c = clampi((int)floorf((x + 1.0f) * S.inv_delta), 0, e - 1);
            int r = clampi((int)floorf((y + 1.0f) * S.inv_delta), 0, e - 1);
            const int sc = d.x > 0.0f ? 1 : (d.x < 0.0f ? -1 : 0);
            const int sr = d.y > 0.0f ? 1 : (d.y < 0.0f ? -1 : 0);
            const float tdx = sc ? S.delta * fabsf(rdx) : INFINITY;
            const float tdy = sr ? S.delta * fabsf(rdy) : INFINITY;
            float tnx = sc ? ((coord(S, c + (sc > 0 ? 1 : 0)) - o.x) * rdx) : INFINITY;
            float tny = sr ? ((coord(S, r + (sr > 0 ? 1 : 0)) - o.y) * rdy) : INFINITY;
            Hit cell_best = miss_hit();
            float tcur = t0;  // where the ray enters the current cell (or starts)
#pragma unroll 1
            for (int guard = 0; guard < 4 * R + 8; ++guard) {
                const float h00 = height(S, r, c), h01 = height(S, r, c + 1);
                const float h10 = height(S, r + 1, c), h11 = height(S, r + 1, c + 1);
                const float texit = fminf(tnx, tny);
                // Cull: a hit on the cell's triangles lies over the cell, i.e.
                // at t in [tcur, texit], and between the cell's lowest and
                // highest vertex; if the ray's z over that interval stays
                // clear of those heights (by 1e-3 relative, far more than rounding) the
                // triangles cannot be hit.  Rays inside the glass, below the
                // surface, skip most cells this way.
                bool test = true;
                if (texit < INFINITY) {
                    const float za = fmaf(tcur, d.z, o.z), zb = d.z != 0.0f ? fmaf(texit, d.z, o.z) : o.z;
                    const float cmin = fminf(fminf(h00, h01), fminf(h10, h11));
                    const float cmax = fmaxf(fmaxf(h00, h01), fmaxf(h10, h11));
                    const float m = kCullZ * (1.0f + fmaxf(fabsf(za), fabsf(zb)));
                    test = fmaxf(za, zb) >= cmin - m && fminf(za, zb) <= cmax + m;
                }
                if (test) {
                    const float x0 = coord(S, c), x1 = coord(S, c + 1), y0 = coord(S, r), y1 = coord(S, r + 1);
                    quad_test(o, d, v3(x0, y0, h00), v3(x1, y0, h01), v3(x0, y1, h10), v3(x1, y1, h11), kTopCode,
                              B.top + 2 * (r * e + c), cell_best);
                }
                if (cell_best.obj >= 0 && cell_best.t <= texit * (1.0f + 1e-6f) + 1e-6f) break;
                if (texit > t1) break;
                tcur = texit;
                if (tnx < tny) {
                    c += sc;
                    tnx += tdx;
                } else {
                    r += sr;
                    tny += tdy;
                }
                if (c < 0 || c >= e || r < 0 || r >= e) break;
            }
            top = cell_best;
        }
    }
    // closest t found so far (any order), and whether a plane parameter is past it
    float tk = top.obj >= 0 ? top.t : INFINITY;
    auto beyond = [&](float tp) { return tp > fmaf(tk, 1e-5f, tk) + 1e-6f; };
    // bottom plane z = 0: the two triangles of the cell under the plane hit
    // (its 3x3 neighbourhood only if rounding put the hit across a cell edge).
    // Every bottom triangle lies in that plane, so its intersection parameter
    // is tp up to rounding (~1e-7 here): with tp < kTMin / 2 none of them can
    // pass the t >= kTMin test -- the case of a ray leaving the bottom face,
    // whose own cell and all 8 neighbours would otherwise be tested in vain.
    Hit bot = miss_hit();
    if (d.z != 0.0f) {
        const float tp = -o.z * rdz;
        if (tp >= 0.5f * kTMin && !beyond(tp)) {
            const float x = fmaf(tp, d.x, o.x), y = fmaf(tp, d.y, o.y);
            const float eps = 1e-5f;
            if (x >= -1.0f - eps && x <= 1.0f + eps && y >= -1.0f - eps && y <= 1.0f + eps) {
                const int c = clampi((int)floorf((x + 1.0f) * S.inv_delta), 0, e - 1);
                const int r = clampi((int)floorf((y + 1.0f) * S.inv_delta), 0, e - 1);
#pragma unroll 1
                for (int j = 0; j < 9; ++j) {
                    // j = 0: the cell itself; 1..8: its neighbours, row-major, if it held no hit
                    if (j == 1 && bot.obj >= 0) break;
                    int rr = r, cc = c;
                    if (j > 0) {
                        const int q = j - 1 + (j > 4 ? 1 : 0);
                        rr = r + q / 3 - 1;
                        cc = c + q % 3 - 1;
                        if (rr < 0 || rr >= e || cc < 0 || cc >= e) continue;
                    }
                    const float x0 = coord(S, cc), x1 = coord(S, cc + 1), y0 = coord(S, rr), y1 = coord(S, rr + 1);
                    quad_test(o, d, v3(x0, y0, 0.f), v3(x1, y0, 0.f), v3(x0, y1, 0.f), v3(x1, y1, 0.f), kBottomCode,
                              B.bottom + 2 * (rr * e + cc), bot);
                }
            }
        }
    }
    if (bot.obj >= 0) tk = fminf(tk, bot.t);
    // ground plane (2 triangles), skipped when it lies behind the ray (as for
    // the bottom plane: both triangles would fail t >= kTMin) or past tk
    Hit best = miss_hit();
    if (d.z != 0.0f) {
        const float tg = (S.sensor_z - o.z) * rdz;
        if (tg >= 0.5f * kTMin && !beyond(tg)) {
            mt_test(o, d, v3(-1.f, -1.f, S.sensor_z), v3(1.f, -1.f, S.sensor_z), v3(-1.f, 1.f, S.sensor_z), 1, 0,
                    best);
            mt_test(o, d, v3(-1.f, 1.f, S.sensor_z), v3(1.f, -1.f, S.sensor_z), v3(1.f, 1.f, S.sensor_z), 1, 1, best);
        }
    }
    // in test order: ground, then bottom
    if (bot.obj >= 0 && (best.obj < 0 || bot.t < best.t)) best = bot;
    if (best.obj >= 0) tk = fminf(tk, best.t);
    // side walls x = -1, x = +1, y = -1, y = +1: the two triangles of the
    // segment at the plane hit, when that hit lies within the wall's extent
    // (the neighbouring segments if they miss).  A wall hit needs the ray to
    // cross |x| = 1 or |y| = 1 at a height within the slab; if the ray's xy
    // stays inside the square at both ends of its slab interval (widened by
    // 1e-3 in z, 1e-4 in xy: far above rounding) it stays inside in between,
    // and the four walls are not tested.
    bool walls = true;
    if (d.z != 0.0f) {
        float ta = (-1e-3f - o.z) * rdz, tb = (S.hmax + 1e-3f - o.z) * rdz;
        if (ta > tb) {
            const float tmp = ta;
            ta = tb;
            tb = tmp;
        }
        ta = fmaxf(ta, 0.0f);
        if (!(ta <= tb)) {
            walls = false;  // the slab lies behind the ray
        } else {
            const float lim = 1.0f - 1e-4f;
            walls = !(fabsf(fmaf(ta, d.x, o.x)) < lim && fabsf(fmaf(ta, d.y, o.y)) < lim &&
                      fabsf(fmaf(tb, d.x, o.x)) < lim && fabsf(fmaf(tb, d.y, o.y)) < lim);
        }
    }
#pragma unroll 1
    for (int w = 0; w < (walls ? 4 : 0); ++w) {
        const bool xw = w < 2;
        const float dd = xw ? d.x : d.y;
        if (dd == 0.0f) continue;
        const float P = (w & 1) ? 1.0f : -1.0f;
        const float tp = (P - (xw ? o.x : o.y)) * (xw ? rdx : rdy);
        const float lat = xw ? fmaf(tp, d.y, o.y) : fmaf(tp, d.x, o.x);
        const float z = fmaf(tp, d.z, o.z);
        if (!(tp >= 0.0f) || !(z >= -1e-5f) || !(z <= S.hmax + 1e-5f) || !(fabsf(lat) <= 1.0f + 1e-5f)) continue;
        if (beyond(tp)) continue;
        const int k = clampi((int)floorf((lat + 1.0f) * S.inv_delta), 0, e - 1);
        const int base = w == 0 ? B.left : (w == 1 ? B.right : (w == 2 ? B.front : B.back));
        const int code = w == 0 ? kWallCode[0] : (w == 1 ? kWallCode[1] : (w == 2 ? kWallCode[2] : kWallCode[3]));
        const int fixed = (w & 1) ? e : 0;  // the wall's grid column (x walls) or row (y walls)
        const Hit keep = best;
#pragma unroll 1
        for (int j = 0; j < 3; ++j) {
            // j = 0: the segment itself; then k - 1 and k + 1 if it held no hit
            if (j == 1 && !same_hit(best, keep)) break;
            const int kk = k + (j == 0 ? 0 : (j == 1 ? -1 : 1));
            if (kk < 0 || kk >= e) continue;
            const int ra = xw ? kk : fixed, ca = xw ? fixed : kk;
            const int rb = xw ? kk + 1 : fixed, cb = xw ? fixed : kk + 1;
            const float xa = coord(S, ca), ya = coord(S, ra), xb = coord(S, cb), yb = coord(S, rb);
            quad_test(o, d, v3(xa, ya, 0.f), v3(xa, ya, height(S, ra, ca)), v3(xb, yb, 0.f),
                      v3(xb, yb, height(S, rb, cb)), code, base + 2 * kk, best);
        }
        if (best.obj >= 0) tk = fminf(tk, best.t);
    }
    // the top surface comes last in test order
    if (top.obj >= 0 && (best.obj < 0 || top.t < best.t)) best = top;
    return best;
}

// ---------------------------------------------------------------------------
// shading and optics
// ---------------------------------------------------------------------------

// model/renderable_object.py:383-389 (cw = False)
__device__ __forceinline__ V3 geometric_normal(const TriVerts& T) {
    return normalize(cross(T.v[1] - T.v[0], T.v[2] - T.v[0]));
}

// model/utils.py:65-74 + renormalisation (model/renderable_object.py:505-506)
__device__ __forceinline__ V3 shading_raw(const TriVerts& T, float b1, float b2) {
    const float w0 = 1.0f - b1 - b2;
    return w0 * T.n[0] + b1 * T.n[1] + b2 * T.n[2];
}

}  // namespace hf
