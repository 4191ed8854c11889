// Host build of the height-field tracer (shapefromcaustics_amd/csrc/hf_device.hpp,
// HIP include stripped, __device__ empty) against brute force: every triangle
// of the scene through the same intersection routine, closest hit kept.
// Random rays start above the glass, inside it, on its bottom face and
// outside its sides, in random, vertical and grazing directions.
// Prints "rays hits mismatches"; exits 1 on any mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#define __device__
#define __forceinline__ inline
// the tracer's hardware reciprocal (1 ulp on the GPU) as a host division: it
// only steers which triangles are tested, never a hit's values
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#include "hf_device_host.hpp"

using namespace hf;

int main(int argc, char** argv) {
    const int rays_per_field = argc > 1 ? atoi(argv[1]) : 4000;
    std::mt19937 g(7);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    long total = 0, hits = 0, bad = 0;
    for (int R : {9, 33}) {
        for (int field = 0; field < 2; ++field) {
            std::vector<float> h(R * R), n(3 * R * R, 0.f);
            const float a1 = 0.05f + 0.1f * U(g), f1 = 2 + 6 * U(g);
            float hm = 0, hn = INFINITY;
            for (int r = 0; r < R; ++r)
                for (int c = 0; c < R; ++c) {
                    const float x = -1 + 2.f * c / (R - 1), y = -1 + 2.f * r / (R - 1);
                    h[r * R + c] = 0.2f + a1 * sinf(f1 * x) * cosf(f1 * y);
                    hm = fmaxf(hm, h[r * R + c]);
                    hn = fminf(hn, h[r * R + c]);
                }
            const SceneArgs S{h.data(), n.data(), R, 2.0f / (R - 1), (R - 1) / 2.0f, -10.f, hm, hn};
            const int nprim = 4 * (R - 1) * (R - 1) + 8 * (R - 1);
            for (int k = 0; k < rays_per_field; ++k) {
                const float ox = -1.3f + 2.6f * U(g), oy = -1.3f + 2.6f * U(g);
                const int kind = k % 4;
                const float oz = kind == 0 ? 0.5f + U(g) : kind == 1 ? 0.01f + 0.18f * U(g) : kind == 2 ? 0.0f
                                                                                                      : U(g) - 0.5f;
                float dx = 2 * U(g) - 1, dy = 2 * U(g) - 1, dz = 2 * U(g) - 1;
                if (k % 7 == 0) dx = dy = 0.f, dz = -1.f;
                if (k % 11 == 0) dz *= 0.05f;
                const V3 o{ox, oy, oz}, d{dx, dy, dz};
                const Hit a = trace(S, o, d);
                Hit b = miss_hit();
                for (int p = 0; p < nprim; ++p) {
                    const TriVerts T = tri_verts(S, 0, p);
                    mt_test(o, d, T.v[0], T.v[1], T.v[2], 0, p, b);
                }
                for (int p = 0; p < 2; ++p) {
                    const TriVerts T = tri_verts(S, 1, p);
                    mt_test(o, d, T.v[0], T.v[1], T.v[2], 1, p, b);
                }
                ++total;
                hits += b.obj >= 0;
                const bool same = a.obj == b.obj && a.prim == b.prim && a.t == b.t && a.b1 == b.b1 && a.b2 == b.b2;
                if (!same && bad++ < 5)
                    fprintf(stderr, "R=%d ray %d: trace (%d, %d, t=%.9g) brute force (%d, %d, t=%.9g)\n", R, k, a.obj,
                            a.prim, a.t, b.obj, b.prim, b.t);
            }
        }
    }
    printf("%ld %ld %ld\n", total, hits, bad);
    return bad ? 1 : 0;
}
