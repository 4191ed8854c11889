/*
 * hf.h -- C ABI of the height-field caustic kernels in libpds_hip.so (gfx950).
 *
 * They replace, for a height-field glass substrate above a ground plane, the
 * reference's OptiX stage and its torch optics around the splat:
 *   model/caustics.py:215-221          get_normal_from_height_field (Sobel normals)
 *   PyOptix/RayTrace.cpp:233-291        trace_rays (closest hit, t >= 1e-3)
 *   PyOptix/kernel/ray_programs.cu:84-101  hit / miss values
 *   model/renderable_object.py:52-78, :249-308  substrate triangles (grid_shuffle order)
 *   model/renderable_object.py:383-401, :437-519  geometric / shading normals, dN
 *   model/photon_differential.py:53-78  advance_path, differential transfer
 *   model/caustics.py:152-212          refract_reflect_differentials
 *   model/caustics.py:224-239          change_of_basis_matrix, splat_photons
 *   model/caustics.py:242-291          compute_recursive_refraction (bounce loop)
 * The splat itself is pds_forward_f32 / pds_backward_f32 (pds.h).
 *
 * Gradient semantics are the reference's: the trace's hit distance,
 * barycentrics and triangle index are constants (OptiX returns plain
 * tensors), so dL/dh flows through the shading normals (Sobel of h), the
 * geometric normals (vertex heights) and the differentials, not through the
 * hit positions' dependence on h.  The reference's dN (renderable_object.py
 * :391-401) projects (sum_v L_v . x) n_s orthogonally to n_s and is
 * identically zero; it is zero here.  Paths are capped at HF_MAX_BOUNCES
 * hits (the reference loops until every photon has left the scene,
 * model/caustics.py:255; a GPU thread needs a bound): a photon still
 * bouncing after HF_MAX_BOUNCES hits is dropped, and the hit-recording
 * simulate marks such photons (counted by hf_hits_stats_get).
 *
 * Conventions as in pds.h: device pointers, caller-owned buffers and
 * workspace, a hipStream_t passed as void*, 0 or a negative pds_status.
 */
#ifndef HF_H
#define HF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HF_MAX_CHANNELS 8
#define HF_MAX_BOUNCES 16

/* Height field h [R, R] (row = y, column = x) over [-1, 1]^2, vertex (r, c)
 * at (-1 + 2c/(R-1), -1 + 2r/(R-1), h[r, c]); flat bottom z = 0; ground plane
 * z = sensor_z over [-1, 1]^2.  Photon i has channel i mod channels and sees
 * the index ior[channel] inside the glass (1 outside).  Limits: 2 <= R <=
 * 16384, 1 <= channels <= HF_MAX_CHANNELS, 1 <= n_side <= 32768 and
 * N = n_side^2 * channels <= 2^31 - 1 (else PDS_ERR_ARGUMENT /
 * PDS_ERR_CAPACITY, and the *_bytes queries return 0). */
typedef struct hf_scene {
    int32_t R;
    int32_t channels;
    float sensor_z;
    float ior[HF_MAX_CHANNELS];
    int32_t reserved[4];
} hf_scene;

/* Photon source; photon i is sample i / channels in channel i mod channels,
 * N = n_side^2 * channels photons.
 * kind 0, parallel: vertical rays (0, 0, -1) from height z0 through the
 *   centres of an n_side x n_side grid over [-1, 1]^2 (sample s: row s /
 *   n_side), Du = 2 pitch e_x, Dv = 2 pitch e_y, Dtheta = Dphi = 0, flux
 *   flux_total / n_side^2 each.
 * kind 1, point light (model/caustics.py:26-70): rays from light[] whose
 *   directions are bilinear slerps, with coefficients (a, b) drawn per sample
 *   from (seed, sample) by SplitMix64, of the directions to the corners of the
 *   square [-1, 1]^2 at height plane_z; Du = Dv = 0; Dtheta, Dphi = Frisvad's
 *   orthonormal basis of the direction (model/utils.py:33-53) times
 *   diff_scale (the reference's 2 smoothing sqrt(solid angle / (pi total)));
 *   flux flux_total per photon. */
typedef struct hf_emitter {
    int32_t n_side;
    float z0;
    float flux_total;
    int32_t kind;
    float light[3];
    float plane_z;
    float diff_scale;
    uint32_t seed;
    int32_t reserved[2];
} hf_emitter;

/* Message of the calling thread's last failed hf_* call. */
const char* hf_last_error(void);

/* Device workspace for hf_trace_f32 / hf_simulate_f32 / hf_adjoint_f32. */
size_t hf_workspace_bytes(int32_t R);

/* Unit Sobel normals [R, R, 3] of h (model/caustics.py:215-221, replicate
 * padding, element size 2/(R-1)). */
int hf_normals_f32(const float* h, int32_t R, float* normals, void* stream);

/* trace_rays for this scene (PyOptix/RayTrace.cpp:233-291): origins, dirs
 * [N, 3] -> t [N], uv [N, 2], object [N], primitive [N] (miss: -2, (-1, -1),
 * -1, -1).  normals: hf_normals_f32 of h. */
int hf_trace_f32(const float* h, const float* normals, const hf_scene* scene, const float* origins,
                 const float* dirs, int64_t N, float* t, float* uv, int64_t* object, int64_t* primitive,
                 void* workspace, size_t workspace_bytes, void* stream);

/* Emit, trace, refract and transfer every photon to the ground plane and
 * write its splat inputs (model/caustics.py:233-239) in the pds.h layout:
 * Ep [N], xp [2, N], Mp [2, 3, N], cp [1, N], radius [N], N = n_side^2 *
 * channels.  Photons that miss, are dropped by the reference's masks, or
 * are still bouncing after HF_MAX_BOUNCES hits get cp = -1 (pds_* skip
 * them). */
int hf_simulate_f32(const float* h, const float* normals, const hf_scene* scene, const hf_emitter* emitter,
                    float* Ep, float* xp, float* Mp, int64_t* cp, float* radius, void* workspace,
                    size_t workspace_bytes, void* stream);

/* dL/dh [R, R] (overwritten) from dL/d(Ep, xp, Mp) of the photons of
 * hf_simulate_f32 (as pds_backward_f32 returns them): every photon's path is
 * recomputed and differentiated in reverse. */
int hf_adjoint_f32(const float* h, const float* normals, const hf_scene* scene, const hf_emitter* emitter,
                   const float* grad_Ep, const float* grad_xp, const float* grad_Mp, float* grad_h,
                   void* workspace, size_t workspace_bytes, void* stream);

/* The same pair sharing the forward's ray hits (no reference counterpart; the
 * reference's backward re-runs its OptiX trace through torch autograd,
 * model/caustics.py:242-291): hf_simulate_hits_f32 also records, for every
 * photon that reaches the sensor, its first 4 hits (t, barycentrics,
 * object, primitive) in `hits` (hf_hits_bytes), and hf_adjoint_hits_f32
 * replays those paths without tracing them again (paths with more bounces
 * are traced again).  The adjoint takes `hits` from a simulate on the same
 * h, scene and emitter: the buffer also carries that field's maximum height,
 * which the re-traced paths use.  Every photon's gradient terms are those of
 * hf_adjoint_f32 bit for bit; dL/dh differs only by the (run-dependent)
 * order of the float atomic sums. */
size_t hf_hits_bytes(const hf_scene* scene, const hf_emitter* emitter);

/* How the paths of the hf_simulate_hits_f32 that wrote `hits` ended,
 * counted from the per-photon records on the device (writes only the
 * buffer's counter words; synchronises `stream`: a diagnostic, not for the
 * hot loop).  reached + missed + degenerate + capped = N.  No reference
 * counterpart (the reference drops such photons silently). */
typedef struct hf_path_stats {
    int64_t reached;     /* reached the sensor with a valid splat (cp >= 0) */
    int64_t missed;      /* left the scene: a trace found no hit */
    int64_t degenerate;  /* reached the sensor, dropped by the reference's den != 0 mask */
    int64_t capped;      /* still bouncing after HF_MAX_BOUNCES hits: dropped */
    int64_t reserved[4];
} hf_path_stats;
int hf_hits_stats_get(void* hits, size_t hits_bytes, const hf_scene* scene, const hf_emitter* emitter,
                      hf_path_stats* out, void* stream);
int hf_simulate_hits_f32(const float* h, const float* normals, const hf_scene* scene, const hf_emitter* emitter,
                         float* Ep, float* xp, float* Mp, int64_t* cp, float* radius, void* workspace,
                         size_t workspace_bytes, void* hits, size_t hits_bytes, void* stream);
int hf_adjoint_hits_f32(const float* h, const float* normals, const hf_scene* scene, const hf_emitter* emitter,
                        const float* grad_Ep, const float* grad_xp, const float* grad_Mp, float* grad_h,
                        void* workspace, size_t workspace_bytes, const void* hits, size_t hits_bytes, void* stream);

/* Residual and loss of the Landweber data term (hyperparameter_helper.py:29-30):
 * resid = img - target, loss[0] = 1/2 sum resid^2 (device scalar; FP32
 * per-block partials summed in a fixed order). */
size_t hf_residual_workspace_bytes(int64_t n);
int hf_residual_f32(const float* img, const float* target, int64_t n, float* resid, float* loss,
                    void* workspace, size_t workspace_bytes, void* stream);

/* One pixel-Landweber update of the height offset d [R, R] in place
 * (shape_from_caustics.py:387-402, default options): d - tau grad, masked
 * extended soft shrinkage to [lb, ub] with threshold tau alpha
 * (:184-189), then the volume heuristic VconsHeur(d1, d2, gamma, volume_radius)
 * (:228-241).  mask [R, R] is 0/1. */
typedef struct hf_landweber {
    float tau, alpha, lower, upper, d1, d2, gamma;
    int32_t volume_radius;
    int32_t reserved[4];
} hf_landweber;
size_t hf_landweber_workspace_bytes(int32_t R);
int hf_landweber_f32(float* d, const float* grad, const float* mask, int32_t R, const hf_landweber* p,
                     void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HF_H */
