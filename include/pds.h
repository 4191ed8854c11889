/*
 * pds.h -- C ABI of the MI355X (gfx950) photon-differential splatting library
 * (libpds_hip.so).
 *
 * The library replaces the reference's CUDA extension ``PhotonDifferentialSplatting``:
 *   PyOptix/PhotonDifferentialSplattig.cpp:12-25   pds_forward  (checks + launch)
 *   PyOptix/PhotonDifferentialSplattig.cpp:27-32   pds_backward (checks + launch)
 *   PyOptix/kernel/photon_differentials.cu:115-134 pds_cuda_forward  (alloc + kernel)
 *   PyOptix/kernel/photon_differentials.cu:223-247 pds_cuda_backward (alloc + kernel)
 * Only plain pointers, sizes and an opaque stream cross this boundary; the
 * Python binding (shapefromcaustics_amd/PhotonDifferentialSplatting.py) keeps
 * the reference's tensor-level signatures on top of it.
 *
 * Conventions
 *  - All arrays are device pointers in the reference's layout, contiguous:
 *      Ep [N] f32, xp [2,N] f32, Mp [2,3,N] f32, cp [1,N] i64, radius [N] f32,
 *      image / grad [C,H,W] f32, grad_Ep [N], grad_xp [2,N], grad_Mp [2,3,N].
 *  - Buffers are caller-owned.  The forward zero-fills ``out`` itself.
 *  - ``stream`` is a hipStream_t (NULL = the null stream); every launch is
 *    stream-ordered on it and no call synchronises the host.
 *  - ``workspace`` is caller-owned device scratch of at least
 *    pds_forward_workspace_bytes(...) / pds_backward_workspace_bytes(...)
 *    bytes (it may be NULL when that size is 0).  Its contents need no
 *    initialisation and are meaningless after the call.
 *  - Return value: PDS_OK (0) or a negative pds_status; no exception crosses
 *    the ABI.  pds_last_error() gives the message of the calling thread's
 *    last failure.
 *  - Deliberate supersets of the reference, all documented: negative channels
 *    are skipped (the reference indexes out of bounds), skipped photons get
 *    zero gradients (the reference leaves them uninitialised), the r = 0
 *    gradient is its finite limit (the reference divides 0/0), photon counts
 *    are 64-bit (the reference uses 32-bit accessors).
 */
#ifndef PDS_H
#define PDS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pds_status {
    PDS_OK = 0,
    PDS_ERR_ARGUMENT = -1,    /* invalid sizes, null pointers, bad mode */
    PDS_ERR_WORKSPACE = -2,   /* workspace smaller than the pds_*_workspace_bytes() size */
    PDS_ERR_HIP = -3,         /* a HIP runtime call or launch failed */
    PDS_ERR_CAPACITY = -4     /* a size beyond a documented maximum */
} pds_status;

typedef enum pds_mode {
    PDS_MODE_AUTO = 0,          /* fastest mode: tile-binned when the problem can be binned, else photon order */
    PDS_MODE_PLAIN = 1,         /* one thread per photon, global atomics: the reference's schedule */
    PDS_MODE_BINNED = 2,        /* tile-binned LDS forward, atomic flush (run-to-run order may vary) */
    PDS_MODE_DETERMINISTIC = 3  /* tile-ordered forward, bitwise reproducible run to run */
} pds_mode;

/* Tuning knobs of the binned path; zero fields select the built-in default.
 * None of them changes results beyond FP32 summation order: the semantic
 * window (max_pixel_radius) is never altered. */
typedef struct pds_tuning {
    int32_t tile;          /* square screen tile edge in pixels: 8, 16, 32 or 64 */
    int32_t seg_photons;   /* photon records per tile segment (one workgroup each), <= 2048 */
    int32_t block_photons; /* photons per binning workgroup (default 8192; for fewer than
                              2^21 photons N / 256, but at least 1024) */
    int32_t capacity_pct;  /* records reserved per 100 photons (default 200); more
                              records than that run the photon-order path (so do
                              capacities of 2^31 records or more, which the
                              binned path's 32-bit record offsets cannot hold) */
    int32_t walk;          /* 0: walk each footprint's whole bounding box (default);
                              1: walk only the ellipse's row chords (same results,
                              bit for bit: the pixels left out contribute exactly
                              zero; fewer pixels, more setup per row) */
    int32_t reserved[3];
} pds_tuning;

/* Bytes of device workspace pds_forward_f32 needs for this problem in
 * ``mode`` with tuning ``t`` (NULL = defaults); 0 if it needs none. */
size_t pds_forward_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                                   int32_t mode, const pds_tuning* t);

/* Bytes of device workspace pds_backward_f32 needs (0 if none). */
size_t pds_backward_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius,
                                    int32_t mode, const pds_tuning* t);

/* Forward splat (reference pds_forward, PhotonDifferentialSplattig.cpp:12-25;
 * kernel photon_differentials.cu:60-113).  out [C,H,W] is overwritten. */
int pds_forward_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                    const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                    int32_t max_pixel_radius, int32_t mode, const pds_tuning* t, float* out,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Adjoint gather (reference pds_backward, PhotonDifferentialSplattig.cpp:27-32;
 * kernel photon_differentials.cu:136-221).  Writes all of grad_Ep [N],
 * grad_xp [2,N] and grad_Mp [2,3,N] (grad_Mp[:,2] = 0). */
int pds_backward_f32(const float* grad, const float* Ep, const float* xp, const float* Mp,
                     const int64_t* cp, const float* radius, int64_t N, int32_t C, int32_t H,
                     int32_t W, int32_t max_pixel_radius, int32_t mode, const pds_tuning* t,
                     float* grad_Ep, float* grad_xp, float* grad_Mp, void* workspace,
                     size_t workspace_bytes, void* stream);

/* ---- forward without a workspace ----------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pds_forward_f32 with the same arguments and results,
 * whose scratch is instead a fixed device array of PDS_STATIC_SCRATCH_BYTES
 * inside the library.  Each GPU the library runs on holds its own array
 * until the process ends.  A problem whose pds_forward_workspace_bytes
 * exceeds that size returns PDS_ERR_CAPACITY (2^22 photons into 3 x 1024^2
 * fit with capacity_pct <= 190).  The call allocates nothing.  Every call on
 * one GPU shares that GPU's array, which keeps the last call's contents
 * (each call initialises what it reads): calls that overlap on one GPU --
 * on other streams or from other host threads -- race, and the caller must
 * serialise them. */
#define PDS_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int pds_forward_noalloc_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                            const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                            int32_t max_pixel_radius, int32_t mode, const pds_tuning* t, float* out,
                            void* stream);

/* ---- forward/backward pair sharing one binning -------------------------
 * The autograd path runs the forward and then the backward of the same
 * photons.  pds_forward_bins_f32 leaves the photons' tile binning in the
 * caller-owned ``bins`` buffer and pds_backward_bins_f32 reuses it instead of
 * binning again.  The inputs must be unchanged between the two calls, and so
 * must N, C, H, W, max_pixel_radius, mode and the tuning: they fix the
 * layout of ``bins`` (its segments, records and kept walk order), which the
 * backward does not check.  No counterpart in the reference (its backward
 * re-walks every window). */

/* Bytes of ``bins`` for this problem; 0 if it cannot be binned (tiny image,
 * max_pixel_radius > 63, too many tiles, plain mode): use the plain pair. */
size_t pds_bins_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t max_pixel_radius, int32_t mode,
                      const pds_tuning* t);

/* pds_forward_f32 that keeps the binning in ``bins``. */
int pds_forward_bins_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                         const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                         int32_t max_pixel_radius, int32_t mode, const pds_tuning* t, float* out,
                         void* bins, size_t bins_bytes, void* stream);

/* What a binning in ``bins`` holds, for tuning: filled by pds_bins_stats()
 * after pds_forward_bins_f32 on the same arguments.  No counterpart in the
 * reference. */
typedef struct pds_bins_stats {
    int64_t records;      /* (photon, tile) records written */
    int64_t capacity;     /* records the workspace holds */
    int32_t segments;     /* tile segments (one workgroup each) */
    int32_t overflow;     /* 1: records exceeded capacity, the photon-order path ran */
    int32_t tile;         /* tile edge used */
    int32_t seg_photons;  /* records per segment used */
    int32_t bins;         /* channel x tile bins */
    int32_t block_photons;/* photons per binning workgroup used */
    int32_t reserved[2];
} pds_bins_stats;

/* Read the statistics of ``bins``; synchronises ``stream`` (diagnostic, not
 * for the hot loop). */
int pds_bins_stats_get(const void* bins, size_t bins_bytes, int64_t N, int32_t C, int32_t H, int32_t W,
                       int32_t max_pixel_radius, int32_t mode, const pds_tuning* t, pds_bins_stats* out,
                       void* stream);

/* pds_backward_f32 reusing ``bins`` from pds_forward_bins_f32 on the same inputs
 * and arguments (see above). */
int pds_backward_bins_f32(const float* grad, const float* Ep, const float* xp, const float* Mp,
                          const int64_t* cp, const float* radius, int64_t N, int32_t C, int32_t H,
                          int32_t W, int32_t max_pixel_radius, int32_t mode, const pds_tuning* t,
                          float* grad_Ep, float* grad_xp, float* grad_Mp, void* bins, size_t bins_bytes,
                          void* stream);

/* ---- batches with a device-side photon count ---------------------------
 * No counterpart in the reference (it splats its num_simulations inner
 * simulations one after another, shape_from_caustics.py:351-353).  The
 * bins-sharing pair for a batch buffer of capacity N whose photons beyond a
 * count held in device memory are unused: only photons i < *live_count
 * (clamped to [0, N]; read by the kernels when they run, so one captured HIP
 * graph serves every count) are splatted and get gradients, and the
 * gradients of photons i >= *live_count are not written.  N sizes the bins
 * (pds_bins_bytes(N, ...)); the work follows the live count.  The backward
 * takes the same live_count as its forward. */
int pds_forward_bins_live_f32(const float* Ep, const float* xp, const float* Mp, const int64_t* cp,
                              const float* radius, int64_t N, int32_t C, int32_t H, int32_t W,
                              int32_t max_pixel_radius, int32_t mode, const pds_tuning* t,
                              const int64_t* live_count, float* out, void* bins, size_t bins_bytes,
                              void* stream);
int pds_backward_bins_live_f32(const float* grad, const float* Ep, const float* xp, const float* Mp,
                               const int64_t* cp, const float* radius, int64_t N, int32_t C, int32_t H,
                               int32_t W, int32_t max_pixel_radius, int32_t mode, const pds_tuning* t,
                               const int64_t* live_count, float* grad_Ep, float* grad_xp, float* grad_Mp,
                               void* bins, size_t bins_bytes, void* stream);

/* Message of the calling thread's last failed call ("" if none). */
const char* pds_last_error(void);

/* Library version string. */
const char* pds_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PDS_H */
