/*
 * pds_gen.h -- counter-based photon source of libpds_hip.so (gfx950).
 *
 * No counterpart in the reference: its photons come from th.rand
 * (model/caustics.py:27) and are kept in memory for the backward.  Here
 * photon i of view v is a pure function of (seed, v, i) (SplitMix64 draws),
 * so a backward pass regenerates exactly the photons of its forward pass
 * instead of storing them: the recomputation checkpoint of the scale-out
 * workload (4 views x 2^24 photons -> 12 x 2048 x 2048).
 */
#ifndef PDS_GEN_H
#define PDS_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the calling thread's last failed pds_gen_* call. */
const char* pds_gen_last_error(void);

/* Caustic-line photons of one view into destination indices
 * [offset, offset + n) of arrays laid out as pds.h's with N = stride:
 * 80 % on n_lines thin Gaussian lines (lines [n_lines, 4] on the device:
 * centre x, y in pixels, axis angle, long sigma in pixels; short sigma =
 * long / 20), 20 % uniform on the W x H image; footprint axes
 * s_major ~ logU(0.3, 6) px, condition ~ U(1, 10), orientation U(0, pi);
 * energy U(0.5, 1.5) / n; channel 3 view + i mod 3. */
int pds_gen_caustic_f32(uint64_t seed, int32_t view, int64_t n, int32_t W, int32_t H, const float* lines,
                        int32_t n_lines, float* Ep, float* xp, float* Mp, int64_t* cp, float* radius,
                        int64_t stride, int64_t offset, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PDS_GEN_H */
