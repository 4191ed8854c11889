// pds_binned.hip -- tile-binned forward splat for MI355X.
//
// The reference scatters every photon with global atomics
// (PyOptix/kernel/photon_differentials.cu:60-113).  Global float atomics from
// 256 CUs onto caustic hot spots serialise, so this forward instead sorts the
// photons' (channel, screen tile) overlaps into bins and accumulates each tile
// in LDS, flushing it to HBM with coalesced stores:
//
//   0. init      zero the per-bin totals and the control block
//   1. count     per-block LDS histogram of bin overlaps; one device atomic per
//                (block, non-empty bin) on totals[bin] reserves the block's
//                range inside the bin -> hist[G][nbins] = offset within bin
//   2. plan      exclusive scan of totals, segment list (<= seg records each),
//                capacity check, fixed-point scale for the deterministic mode
//                (up to kFusedPlanBins bins: done by every scatter workgroup
//                in its prologue instead of a launch of its own)
//   3. scatter   each block re-walks its photons and writes one 32-byte record
//                per overlap at slot = binoff + atomicAdd_LDS(cursor[bin], 1);
//                then clears the tiles of bins split over several segments
//   4. tile      one workgroup per segment: its records sorted by walk length,
//                LDS accumulation, flush (store, or atomic add when split)
//   5. convert   deterministic mode: int64 fixed point -> float for split bins
//
// Atomic mode accumulates in a double LDS tile (summation order varies run to
// run; the flush rounds each pixel to float once).
// Deterministic mode accumulates every contribution as a 64-bit fixed-point
// integer (quantum 2^e with 2^e * 2^61 >= N max(E) 3/pi and 2^e * 2^51 >=
// max(E) 3/pi, far below FP32 resolution of any pixel), so the sum is exact
// and order-independent: the output is bitwise reproducible.  +inf
// contributions set bit 62.
//
// If the records exceed the workspace capacity, the same launches switch to
// the photon-order path (scatter zero-fills, tile runs the flattened walk with
// global atomics / global int64 atomics, convert converts the whole image);
// a device flag written by the plan chooses, so the host never synchronises.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "pds_binned.hpp"

namespace pds {

namespace {

constexpr int kCountThreads = 512;    // threads of a binning workgroup: count pass
constexpr int kScatterThreads = 512;  // ... and scatter pass
// The binning workgroups sweep the photon arrays together: round r gives
// workgroup b the kChunk photons of chunk r G + b.  At any moment the G
// workgroups then read one window of G kChunk photons of every input array,
// not G places spread over all of them: the count and scatter passes (each
// reading ~8 arrays) took 148 -> 113 us together with this order instead of
// one contiguous range per workgroup, on the benchmark splat.
constexpr int kChunk = 512;
static_assert(kCountThreads == kChunk && kScatterThreads == kChunk, "one photon per thread and chunk");
constexpr int kMaxBins = 16384;
constexpr int kMaxBinBlocks = 1024;
constexpr int kTileThreads = 512;
constexpr int kMaxSeg = 2048;  // records per segment (sizes the LDS sort of a segment)
// The tile adjoint stages the partials of one part of kStageRecs of a
// segment's records (by storage position) at a time in LDS (32 KB) and stores
// them coalesced.  (Parts of 512 records, 16 KB, which let four workgroups
// share a CU instead of three, measured slower: backward 176 against 160 us,
// with twice the workgroup barriers and walk-order groups half the size.)
constexpr int kStageRecs = 1024;
constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v >> 1); }
constexpr int kStageParts = kMaxSeg / kStageRecs;
static_assert(kStageParts >= 1 && kStageParts <= 4 && kMaxSeg % kStageRecs == 0, "at most 4 parts (16-bit packed counts)");
constexpr int kSortKeys = 256;  // walk lengths 0..255 (clamped), sorted by one LDS counting sort
constexpr int kMinBins = 256;
// Up to this many bins, each scatter workgroup plans the segments itself (a
// scan of the bin totals, a few microseconds of its prologue) instead of a
// separate one-workgroup plan launch.
constexpr int kFusedPlanBins = 1024;
constexpr int64_t kDefaultBlockPhotons = 8192;     // photons per binning workgroup
constexpr int64_t kMinDefaultBlockPhotons = 1024;  // ... at least, for small problems
constexpr int64_t kTargetBinBlocks = 256;          // binning workgroups a small problem still gets (one per CU)
constexpr int kDefaultSeg = 2048;
constexpr double kDefaultCapacity = 2.0;  // records per photon reserved in the workspace
constexpr int kMaxRecsPerPhoton = 4;      // backward: photons with more records are recomputed directly
constexpr uint8_t kSpecial = 255;         // class of photons the backward recomputes directly

thread_local const char* g_err = "";

struct Ctrl {
    uint32_t nseg;        // segments planned
    uint32_t overflow;    // 1: records exceed capacity -> photon-order path
    uint32_t max_e_bits;  // max finite positive energy (float bits), deterministic mode
    int32_t fx_exp;       // fixed-point quantum exponent e (quantum 2^e)
    uint64_t total;       // records
};

struct Geom {
    int32_t T, tshift, tiles_x, tiles_y, nbins, G, seg;
    int64_t per_block;  // photons per binning workgroup (at most; chunk-interleaved)
    int64_t rounds;     // chunks per binning workgroup: ceil(N / (G kChunk))
    int64_t cap_records, max_segments;
    bool det;
    bool box_walk;    // walk whole footprint boxes (default) instead of the ellipse's row chords
    bool fused_plan;  // the scatter plans the segments (no plan launch)
};

struct Layout {
    size_t hist, totals, binoff, emax, segs, ctrl, recs, slots, info, partials, order, acc64, end;
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int ilog2(int v) {
    int s = 0;
    while ((1 << s) < v) ++s;
    return s;
}

// Default tile edge: the largest of 64, 32, 16 that still gives at least
// kMinBins bins (a few per CU pair), since larger tiles amortise the flush and
// keep more of a footprint in one tile; 8 for images too small for that.
int default_tile(int32_t C, int32_t H, int32_t W) {
    for (int T = 64; T >= 16; T >>= 1) {
        const int64_t bins = (int64_t)C * ((W + T - 1) / T) * ((H + T - 1) / T);
        if (bins >= kMinBins) return T;
    }
    return 8;
}

// Tile edge: the tuned one if given, else the default, raised until the bins
// fit the LDS histogram.  false: the binned path cannot run (photon order).
bool make_geom(int64_t N, int32_t C, int32_t H, int32_t W, int32_t cap, int32_t mode, const pds_tuning* t, Geom& g) {
    if (cap > 63) return false;  // packed record holds window radii < 64
    int T = (t && t->tile > 0) ? t->tile : default_tile(C, H, W);
    if (T != 8 && T != 16 && T != 32 && T != 64) return false;
    for (;;) {
        const int64_t tx = (W + T - 1) / T, ty = (H + T - 1) / T;
        if ((int64_t)C * tx * ty <= kMaxBins) break;
        if (T >= 64) return false;
        T *= 2;
    }
    g.T = T;
    g.tshift = ilog2(T);
    g.tiles_x = (W + T - 1) / T;
    g.tiles_y = (H + T - 1) / T;
    g.nbins = C * g.tiles_x * g.tiles_y;
    g.seg = (t && t->seg_photons > 0) ? t->seg_photons : kDefaultSeg;
    if (g.seg > kMaxSeg) g.seg = kMaxSeg;
    // default photons per binning block: 8192, fewer for small problems so that
    // count and scatter still launch ~256 workgroups, but at least 1024
    const int64_t bp_default = std::min(kDefaultBlockPhotons, std::max(kMinDefaultBlockPhotons,
                                                                        (N + kTargetBinBlocks - 1) / kTargetBinBlocks));
    const int64_t bp = (t && t->block_photons > 0) ? (int64_t)t->block_photons : bp_default;
    int64_t G = (N + bp - 1) / bp;
    if (G > kMaxBinBlocks) G = kMaxBinBlocks;
    if (G < 1) G = 1;
    g.G = (int32_t)G;
    g.per_block = (N + G - 1) / G;
    g.rounds = std::max((int64_t)1, (N + G * kChunk - 1) / (G * kChunk));
    double capf = kDefaultCapacity;
    if (t && t->capacity_pct > 0) capf = t->capacity_pct / 100.0;
    g.cap_records = (int64_t)(capf * (double)N) + 1024;
    // record offsets and segment starts are kept in 32-bit signed fields
    if (g.cap_records >= ((int64_t)1 << 31)) return false;
    g.max_segments = g.nbins + g.cap_records / g.seg + 1;
    g.det = mode == PDS_MODE_DETERMINISTIC;
    g.box_walk = !(t && t->walk == 1);
    g.fused_plan = g.nbins <= kFusedPlanBins;
    return true;
}

// keep: also the per-photon record slots and classes and the per-record
// partial gradients the backward needs.
Layout make_layout(const Geom& g, int64_t N, int32_t C, int32_t H, int32_t W, bool keep) {
    Layout L;
    size_t o = 0;
    L.hist = o;
    o = align256(o + sizeof(uint32_t) * (size_t)g.G * g.nbins);
    L.totals = o;
    o = align256(o + sizeof(uint32_t) * (size_t)g.nbins);
    L.binoff = o;
    o = align256(o + sizeof(uint32_t) * (size_t)g.nbins);
    L.emax = o;
    o = align256(o + sizeof(float) * (size_t)g.G);
    L.segs = o;
    o = align256(o + sizeof(int4) * (size_t)g.max_segments);
    L.ctrl = o;
    o = align256(o + sizeof(Ctrl));
    L.recs = o;
    o = align256(o + 32 * (size_t)g.cap_records);
    L.slots = L.info = L.partials = L.order = o;
    if (keep) {
        o = align256(o + sizeof(uint32_t) * kMaxRecsPerPhoton * (size_t)N);
        L.info = o;
        o = align256(o + (size_t)N);
        L.partials = o;
        o = align256(o + 32 * (size_t)g.cap_records);
        L.order = o;  // the forward's walk order of every segment, for the backward
        o = align256(o + sizeof(uint32_t) * (size_t)g.cap_records);
    }
    L.acc64 = o;
    if (g.det) o = align256(o + sizeof(uint64_t) * (size_t)C * H * W);
    L.end = o;
    return L;
}

// ---------------------------------------------------------------------------
// photon -> clipped footprint box (forward semantics)
// ---------------------------------------------------------------------------
struct BinPhoton {
    bool rec;      // has records: finite frame and a non-empty footprint box
    bool special;  // the backward recomputes it directly: non-finite energy or frame
    int32_t c;
    int32_t x0, x1, y0, y1;  // box on the image, inclusive
    Frame f;
    float e;
};

// Photon i's footprint box: window (photon_differentials.cu:80-96) n image n
// ellipse bounding box.  Photons outside the channel range or with an empty
// window contribute nothing and have zero gradients (neither rec nor special).
// A non-finite centre, matrix or M[:,2] makes q NaN or infinite, so K = 0 and
// the forward gets nothing from it, but the reference's backward turns it
// into NaN gradients: such photons, and those with a non-finite energy, are
// "special" and the backward evaluates them by a direct serial walk.  Photons
// with E <= 0 get records (their gE is sum K g); the forward's val > 0 test
// (:106) skips their pixels.
// The inputs of one photon, loaded together so that the next photon's loads
// can be in flight while the current one is binned.
struct RawPhoton {
    int64_t c;
    float e, cx, cy, radius, m00, m01, m02, m10, m11, m12;
};

// Streamed data: arrays a kernel reads once, in order, and nothing later in
// the pass reads again (the photons in the count and scatter passes, the
// finalize's classes and slots) are loaded non-temporally, and the classes
// and slots, which only the finalize reads, at the end of the backward, are
// stored so too, as are the finalize's gradients (written once, read by no
// later pass), so that they do not displace from the caches the lines that
// the passes' scattered writes are still filling (record runs, partials).
// Measured on the benchmark splat: forward 229 -> 217 us with the photons,
// backward about 3 and 4 us less with the classes' and slots' loads and
// stores, the next forward 5-7 us shorter with the gradients' stores;
// non-temporal loads of gathered data (records, partials) and non-temporal
// stores of data read again soon (the walk order, the partials) were slower.
template <class T>
__device__ __forceinline__ T stream_load(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void stream_store(T v, T* p) {
    __builtin_nontemporal_store(v, p);
}

// with_e = false: the energy is not read (0); with_m2 = false: nor is M[:, 2]
// (0) -- for passes that only need the photon's footprint, which depends on
// neither
__device__ __forceinline__ RawPhoton load_photon(const PhotonArgs& P, int64_t i, bool with_e = true,
                                                 bool with_m2 = true) {
    const int64_t N = P.N;
    RawPhoton r;
    r.c = stream_load(&P.cp[i]);
    r.e = with_e ? stream_load(&P.E[i]) : 0.0f;
    r.cx = stream_load(&P.xp[i]);
    r.cy = stream_load(&P.xp[N + i]);
    r.radius = stream_load(&P.radius[i]);
    r.m00 = stream_load(&P.M[i]);
    r.m01 = stream_load(&P.M[N + i]);
    r.m02 = with_m2 ? stream_load(&P.M[2 * N + i]) : 0.0f;
    r.m10 = stream_load(&P.M[3 * N + i]);
    r.m11 = stream_load(&P.M[4 * N + i]);
    r.m12 = with_m2 ? stream_load(&P.M[5 * N + i]) : 0.0f;
    return r;
}

__device__ __forceinline__ BinPhoton bin_photon(const RawPhoton& R, const ImageArgs& I) {
    BinPhoton b;
    b.rec = false;
    b.special = false;
    const int64_t c = R.c;
    if (c < 0 || c >= I.C) return b;
    const float e = R.e;
    const float cx = R.cx, cy = R.cy;
    const Window win = photon_window(R.radius, cx, cy, I.W, I.H, I.cap);
    if (win.rx < 0 || win.ry < 0) return b;
    int32_t x0 = (int32_t)max((int64_t)win.pxc - win.rx, (int64_t)0);
    int32_t x1 = (int32_t)min((int64_t)win.pxc + win.rx, (int64_t)I.W - 1);
    int32_t y0 = (int32_t)max((int64_t)win.pyc - win.ry, (int64_t)0);
    int32_t y1 = (int32_t)min((int64_t)win.pyc + win.ry, (int64_t)I.H - 1);
    if (x0 > x1 || y0 > y1) return b;
    const float m00 = R.m00, m01 = R.m01, m10 = R.m10, m11 = R.m11;
    // 0, or NaN when M[:, 2] is not finite: such a photon keeps its records
    // (so the count pass need not read M[:, 2]) with a NaN energy, which the
    // forward's val > 0 test drops, as the reference's NaN q does
    const float poison = R.m02 * 0.0f + R.m12 * 0.0f;
    const bool finite_frame = isfinite(cx) && isfinite(cy) && isfinite(m00) && isfinite(m01) && isfinite(m10) &&
                              isfinite(m11);
    b.special = !finite_frame || !isfinite(e) || poison != 0.0f;
    if (!finite_frame) return b;
    b.f = photon_frame(cx, cy, m00, m01, m10, m11, 0.0f, I.W, I.H, I.inv_hw, I.inv_hh);
    clip_to_ellipse(b.f, x0, x1, y0, y1);
    if (x0 > x1 || y0 > y1) return b;
    b.x0 = x0;
    b.x1 = x1;
    b.y0 = y0;
    b.y1 = y1;
    b.c = (int32_t)c;
    b.e = e + poison;
    b.rec = true;
    return b;
}

// ---------------------------------------------------------------------------
// 32-byte record of one (photon, tile) overlap, with the photon's footprint
// box already clipped to the tile (the scatter has the box; the tile kernels
// then neither recompute the window nor clip to the ellipse):
//   x  fx quantised to 2^-25 px (signed, bits 6-31) | box x origin in the tile (0-5)
//   y  fy likewise                                  | box y origin in the tile (0-5)
//   z  energy (float bits)
//   w  box width - 1 (0-5) | box height - 1 (6-11) | box x origin - ix + 64 (12-19) |
//      box y origin - iy + 64 (20-27)
//   second half: a00, a01, a10, a11 (float bits)
// fx, fy lie in [-1/2, 1/2], so the quantised fraction (|fxq| <= 2^24) is an
// exact float, equal to the float fraction where |fx| >= 1/4 (whose spacing
// is 2^-25 there) and within 1.5e-8 px of it below; box origins relative to
// (ix, iy) lie in [-64, 64] (window radii <= 63).
// ---------------------------------------------------------------------------
constexpr float kFracScale = 33554432.0f;        // 2^25
constexpr float kFracUnit = 1.0f / 33554432.0f;  // 2^-25

__device__ __forceinline__ uint32_t quantise_frac(float f) { return (uint32_t)__float2int_rn(f * kFracScale); }

__device__ __forceinline__ float frac_of(uint32_t word) { return (float)((int32_t)word >> 6) * kFracUnit; }

// Record words x, y, w for the photon box [x0, x1] x [y0, y1] (already within
// the image) in the tile with origin (X0, Y0), edge T; fxq, fyq from quantise_frac.
__device__ __forceinline__ uint4 rec_head(uint32_t fxq, uint32_t fyq, float e, int32_t ix, int32_t iy, int32_t x0,
                                          int32_t x1, int32_t y0, int32_t y1, int32_t X0, int32_t Y0, int32_t T) {
    const int32_t bx0 = max(x0, X0), bx1 = min(x1, X0 + T - 1);
    const int32_t by0 = max(y0, Y0), by1 = min(y1, Y0 + T - 1);
    uint4 q;
    q.x = (fxq << 6) | (uint32_t)(bx0 - X0);
    q.y = (fyq << 6) | (uint32_t)(by0 - Y0);
    q.z = __float_as_uint(e);
    q.w = (uint32_t)(bx1 - bx0) | ((uint32_t)(by1 - by0) << 6) | ((uint32_t)(bx0 - ix + 64) << 12) |
          ((uint32_t)(by0 - iy + 64) << 20);
    return q;
}

// the box walk's length of a record, from its w word: box rows x pixel pairs per row
__device__ __forceinline__ int rec_pairs(uint32_t w) {
    return (int)(((w & 63) + 2) >> 1) * (int)(((w >> 6) & 63) + 1);
}

// ---------------------------------------------------------------------------
// 0. init: totals and the control block (no memset, so the sequence can be
//    captured in a graph)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void bin_init_kernel(Geom g, uint32_t* __restrict__ totals, Ctrl* ctrl) {
    for (int k = threadIdx.x; k < g.nbins; k += 1024) totals[k] = 0u;
    if (threadIdx.x == 0) {
        ctrl->nseg = 0;
        ctrl->overflow = 0;
        ctrl->max_e_bits = 0;
        ctrl->fx_exp = 0;
        ctrl->total = 0;
    }
}

// ---------------------------------------------------------------------------
// 1. count: LDS histogram of the block's overlaps, then one device atomic per
//    non-empty bin reserves the block's range inside that bin (the ranges'
//    order among blocks varies run to run; nothing downstream depends on it:
//    the atomic mode's sums are unordered anyway, the deterministic mode's
//    fixed-point sums are exact, and every record's gradient partial is
//    computed on its own).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kCountThreads) void bin_count_kernel(PhotonArgs P, ImageArgs I, Geom g,
                                                                uint32_t* __restrict__ hist,
                                                                uint32_t* __restrict__ totals,
                                                                float* __restrict__ emax_blk) {
    extern __shared__ uint32_t lh[];  // nbins entries
    __shared__ float wmax[kCountThreads / 64];
    for (int k = threadIdx.x; k < g.nbins; k += kCountThreads) lh[k] = 0;
    __syncthreads();
    // this workgroup's chunk of every round (kChunk)
    const int64_t hi = live_photons(P);
    const int64_t step = (int64_t)g.G * kChunk;
    float emax = 0.0f;
    int64_t i = (int64_t)blockIdx.x * kChunk + threadIdx.x;
    // two photons of this lane in flight while one is binned
    RawPhoton r0 = {}, r1 = {};
    if (i < hi) r0 = load_photon(P, i, g.det, false);
    if (i + step < hi) r1 = load_photon(P, i + step, g.det, false);
    for (; i < hi; i += step) {
        const RawPhoton now = r0;
        r0 = r1;
        if (i + 2 * step < hi) r1 = load_photon(P, i + 2 * step, g.det, false);
        const BinPhoton b = bin_photon(now, I);
        if (!b.rec) continue;
        if (b.e > 0.0f && isfinite(b.e)) emax = fmaxf(emax, b.e);
        const int tx0 = b.x0 >> g.tshift, tx1 = b.x1 >> g.tshift;
        const int ty0 = b.y0 >> g.tshift, ty1 = b.y1 >> g.tshift;
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&lh[(b.c * g.tiles_y + ty) * g.tiles_x + tx], 1u);
    }
    if (g.det) {
        for (int o = 32; o > 0; o >>= 1) emax = fmaxf(emax, __shfl_xor(emax, o));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = emax;
    }
    __syncthreads();
    uint32_t* row = hist + (size_t)blockIdx.x * g.nbins;
    for (int k = threadIdx.x; k < g.nbins; k += kCountThreads) {
        const uint32_t n = lh[k];
        if (n) row[k] = atomicAdd(&totals[k], n);  // entries of empty bins are never read
    }
    if (g.det && threadIdx.x == 0) {
        float m = 0.0f;
        for (int w = 0; w < kCountThreads / 64; ++w) m = fmaxf(m, wmax[w]);
        emax_blk[blockIdx.x] = m;
    }
}

// Quantum exponent of the fixed point for N photons of maximal finite
// positive energy m (the deterministic mode; see to_fixed): N m 3/pi <= 2^(e+61) and m 3/pi <= 2^(e+51).
__device__ __forceinline__ int fixed_exponent(int64_t N, float m) {
    if (!(m > 0.0f)) return 0;
    const double peak = (double)m * 0.9549296585513720;
    return max((int)ceil(log2((double)N * peak)) - 61, (int)ceil(log2(peak)) - 51);
}

// ---------------------------------------------------------------------------
// 2. plan: exclusive scans of the bins' record totals and segment counts,
//    segments of at most g.seg records (every bin gets one, so that its flush
//    writes its tile), capacity check, fixed-point scale of the deterministic
//    mode.  Run by one workgroup of NT threads: the plan launch, or each
//    scatter workgroup for itself when the bins are few (fused_plan).
// ---------------------------------------------------------------------------
struct PlanLds {
    uint64_t wsum[16];
    uint32_t wseg[16];
    float wmax[16];
    uint64_t base_rec;  // after plan_scan: total records
    uint32_t base_seg;  // ... and segments
};

// visit(bin, off, t, soff, ns): called by the thread holding bin, with its
// first record, records, first segment and segments.
template <int NT, class F>
__device__ __forceinline__ void plan_scan(const Geom& g, const uint32_t* __restrict__ totals,
                                          const float* __restrict__ emax_blk, PlanLds& L, F&& visit) {
    static_assert(NT % 64 == 0 && NT <= 1024, "whole wavefronts, at most 16");
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        L.base_rec = 0;
        L.base_seg = 0;
    }
    if (g.det) {
        // maximum finite positive energy over the binning blocks
        float m = 0.0f;
        for (int k = tid; k < g.G; k += NT) m = fmaxf(m, emax_blk[k]);
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0) L.wmax[w] = m;
    }
    __syncthreads();
    // bins in chunks of NT, one per thread, carried across chunks
    for (int c0 = 0; c0 < g.nbins; c0 += NT) {
        const int bin = c0 + tid;
        const uint32_t t = bin < g.nbins ? totals[bin] : 0u;
        const uint32_t ns = bin < g.nbins ? max(1u, (t + g.seg - 1) / g.seg) : 0u;
        // inclusive wave scans (64-bit records, 32-bit segments)
        uint64_t a = t;
        uint32_t b = ns;
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t ya = __shfl_up(a, o);
            const uint32_t yb = __shfl_up(b, o);
            if (lane >= o) {
                a += ya;
                b += yb;
            }
        }
        if (lane == 63) {
            L.wsum[w] = a;
            L.wseg[w] = b;
        }
        __syncthreads();
        uint64_t pa = L.base_rec;
        uint32_t pb = L.base_seg;
        for (int k = 0; k < w; ++k) {
            pa += L.wsum[k];
            pb += L.wseg[k];
        }
        if (bin < g.nbins) visit(bin, pa + a - t, t, pb + b - ns, ns);
        __syncthreads();
        if (tid == NT - 1) {
            L.base_rec = pa + a;
            L.base_seg = pb + b;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ bool plan_overflow(const Geom& g, const PlanLds& L) {
    return L.base_rec > (uint64_t)g.cap_records || L.base_seg > (uint64_t)g.max_segments;
}

// The bin's segments, when they fit the workspace (else the plan overflows
// and nothing reads them).
__device__ __forceinline__ void write_segments(const Geom& g, int4* __restrict__ segs, int bin, uint64_t off,
                                               uint32_t t, uint32_t soff, uint32_t ns) {
    const bool fits = off + t <= (uint64_t)g.cap_records && soff + ns <= (uint64_t)g.max_segments;
    if (fits)
        for (uint32_t s = 0; s < ns; ++s) {
            const uint32_t st = (uint32_t)off + s * g.seg;
            segs[soff + s] = make_int4(bin, (int)st, (int)min((uint32_t)g.seg, t - s * g.seg),
                                       (ns > 1 ? 1 : 0) | (s == 0 ? 2 : 0));
        }
}

// thread 0 of the planning workgroup, after plan_scan
template <int NT>
__device__ __forceinline__ void write_ctrl(const Geom& g, int64_t N, const PlanLds& L, Ctrl* ctrl) {
    ctrl->total = L.base_rec;
    const bool over = plan_overflow(g, L);
    ctrl->overflow = over ? 1u : 0u;
    ctrl->nseg = over ? 0u : L.base_seg;
    // fixed-point quantum of the deterministic mode
    float m = 0.0f;
    if (g.det)
        for (int k = 0; k < NT / 64; ++k) m = fmaxf(m, L.wmax[k]);
    ctrl->max_e_bits = __float_as_uint(m);
    ctrl->fx_exp = fixed_exponent(N, m);
}

__global__ __launch_bounds__(1024) void bin_plan_kernel(Geom g, int64_t N, const uint32_t* __restrict__ totals,
                                                        const float* __restrict__ emax_blk,
                                                        uint32_t* __restrict__ binoff, int4* __restrict__ segs,
                                                        Ctrl* ctrl) {
    __shared__ PlanLds L;
    plan_scan<1024>(g, totals, emax_blk, L, [&](int bin, uint64_t off, uint32_t t, uint32_t soff, uint32_t ns) {
        binoff[bin] = (uint32_t)off;
        write_segments(g, segs, bin, off, t, soff, ns);
    });
    if (threadIdx.x == 0) write_ctrl<1024>(g, N, L, ctrl);
}

// ---------------------------------------------------------------------------
// 3. scatter.  On overflow it zero-fills the output (or the int64 image) for
//    the photon-order path instead.  Afterwards it clears the tiles of split
//    bins, which the tile kernels then accumulate into atomically.  With
//    fused_plan every workgroup first plans from the bin totals itself
//    (binoff unused): the same scans give each workgroup its cursors, the
//    workgroups write the segments of every gridDim.x-th bin, and workgroup 0
//    the control block, which no workgroup of this launch reads.
// ---------------------------------------------------------------------------
template <bool KEEP>
__global__ __launch_bounds__(kScatterThreads) void bin_scatter_kernel(PhotonArgs P, ImageArgs I, Geom g,
                                                                  const uint32_t* __restrict__ hist,
                                                                  const uint32_t* __restrict__ binoff,
                                                                  const uint32_t* __restrict__ totals,
                                                                  const float* __restrict__ emax_blk,
                                                                  int4* __restrict__ segs,
                                                                  uint4* __restrict__ recs, uint32_t* __restrict__ slots,
                                                                  uint8_t* __restrict__ info, Ctrl* ctrl,
                                                                  float* __restrict__ out,
                                                                  unsigned long long* __restrict__ acc64) {
    extern __shared__ uint32_t cur[];  // nbins entries
    const uint32_t* row = hist + (size_t)blockIdx.x * g.nbins;
    bool overflow;
    if (g.fused_plan) {
        __shared__ PlanLds L;
        plan_scan<kScatterThreads>(g, totals, emax_blk, L,
                                   [&](int bin, uint64_t off, uint32_t t, uint32_t soff, uint32_t ns) {
                                       cur[bin] = (uint32_t)off + row[bin];
                                       if (bin % (int)gridDim.x == (int)blockIdx.x)
                                           write_segments(g, segs, bin, off, t, soff, ns);
                                   });
        if (blockIdx.x == 0 && threadIdx.x == 0) write_ctrl<kScatterThreads>(g, P.N, L, ctrl);
        overflow = plan_overflow(g, L);
    } else {
        overflow = ctrl->overflow != 0;
    }
    // the image to clear: none when binning for a backward (no image output)
    const bool clear = g.det ? acc64 != nullptr : out != nullptr;
    if (overflow) {
        if (!clear) return;
        const int64_t n = (int64_t)I.C * I.H * I.W;
        const int64_t stride = (int64_t)gridDim.x * kScatterThreads;
        for (int64_t k = (int64_t)blockIdx.x * kScatterThreads + threadIdx.x; k < n; k += stride) {
            if (g.det)
                acc64[k] = 0ull;
            else
                out[k] = 0.0f;
        }
        return;
    }
    if (!g.fused_plan) {
        for (int k = threadIdx.x; k < g.nbins; k += kScatterThreads) cur[k] = binoff[k] + row[k];
        __syncthreads();
    }
    // this workgroup's chunk of every round, as in the count pass
    const int64_t hi = live_photons(P);
    const int64_t step = (int64_t)g.G * kChunk, lo = (int64_t)blockIdx.x * kChunk;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // Each round of records is staged in LDS and written by lane pairs, one
    // 32-byte record per pair (lane 2s: first half, lane 2s+1: second half of
    // lane s's record), so that the two halves of a record leave in one
    // contiguous request instead of two scattered 16-byte ones.
    __shared__ uint4 stage[kScatterThreads / 64][2][64];
    __shared__ uint32_t stage_slot[kScatterThreads / 64][64];
    int64_t i = lo + threadIdx.x;
    // two photons of this lane in flight while one is binned
    RawPhoton r0 = {}, r1 = {};
    if (i < hi) r0 = load_photon(P, i);
    if (i + step < hi) r1 = load_photon(P, i + step);
    // wavefront-uniform rounds (the pair writes need the whole wavefront)
    for (int64_t first = lo + (threadIdx.x & ~63); first < hi; first += step, i += step) {
        const bool valid = i < hi;
        const RawPhoton now = r0;
        r0 = r1;
        if (i + 2 * step < hi) r1 = load_photon(P, i + 2 * step);
        BinPhoton b = {};
        if (valid) b = bin_photon(now, I);
        int tx0 = 0, ty0 = 0, nx = 1, nrec = 0;
        if (b.rec) {
            tx0 = b.x0 >> g.tshift;
            ty0 = b.y0 >> g.tshift;
            nx = (b.x1 >> g.tshift) - tx0 + 1;
            nrec = nx * ((b.y1 >> g.tshift) - ty0 + 1);
        }
        int kmax = nrec;
        for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o));
        const uint32_t fxq = quantise_frac(b.f.fx), fyq = quantise_frac(b.f.fy);
        const uint4 h1 = make_uint4(__float_as_uint(b.f.a00), __float_as_uint(b.f.a01), __float_as_uint(b.f.a10),
                                    __float_as_uint(b.f.a11));
        int kx = 0, ky = 0;  // this lane's current tile, relative to (tx0, ty0), row-major
        for (int k = 0; k < kmax; ++k) {
            const bool has = k < nrec;
            if (has) {
                const int tx = tx0 + kx, ty = ty0 + ky;
                const uint32_t slot = atomicAdd(&cur[(b.c * g.tiles_y + ty) * g.tiles_x + tx], 1u);
                stage[w][0][lane] = rec_head(fxq, fyq, b.e, b.f.ix, b.f.iy, b.x0, b.x1, b.y0, b.y1, tx << g.tshift,
                                             ty << g.tshift, g.T);
                stage[w][1][lane] = h1;
                stage_slot[w][lane] = slot;
                if (KEEP && k < kMaxRecsPerPhoton) stream_store(slot, &slots[(size_t)k * P.N + i]);
                if (++kx == nx) {
                    kx = 0;
                    ++ky;
                }
            }
            const uint64_t mask = __ballot(has);
            wave_lds_fence();
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int src = (lane >> 1) + 32 * p, h = lane & 1;
                if ((mask >> src) & 1ull) recs[2 * (size_t)stage_slot[w][src] + h] = stage[w][h][src];
            }
            wave_lds_fence();
        }
        if (KEEP && valid) stream_store((b.special || nrec > kMaxRecsPerPhoton) ? kSpecial : (uint8_t)nrec, &info[i]);
    }
    // clear the tiles of split bins (more records than one segment holds)
    if (!clear) return;
    const int tpc = g.tiles_x * g.tiles_y;
    for (int bin = blockIdx.x; bin < g.nbins; bin += gridDim.x) {
        if (totals[bin] <= (uint32_t)g.seg) continue;
        const int c = bin / tpc, t = bin - c * tpc;
        const int ty = t / g.tiles_x, tx = t - ty * g.tiles_x;
        const int X0 = tx << g.tshift, Y0 = ty << g.tshift;
        for (int k = threadIdx.x; k < g.T * g.T; k += kScatterThreads) {
            const int px = X0 + (k & (g.T - 1)), py = Y0 + (k >> g.tshift);
            if (px < I.W && py < I.H) {
                const int64_t o = ((int64_t)c * I.H + py) * I.W + px;
                if (g.det)
                    acc64[o] = 0ull;
                else
                    out[o] = 0.0f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// tile helpers
// ---------------------------------------------------------------------------
// The deterministic mode's fixed point.  Its quantum 2^e (fixed_exponent)
// keeps every pixel's sum below 2^61 and every single contribution below
// 2^51 quanta.  to_fixed(val) for val > 0 is the nearest multiple of 2^e (ties
// to even): x = val 2^-e is exact in double and below 2^52 - 1, so adding 2^52
// rounds it to an integer held in the low 52 bits of the sum -- a few
// instructions instead of a 64-bit integer conversion.  +inf gives 0; its
// pixel is marked by bit 62 instead (fixed_deposit).
constexpr unsigned long long kInfMark = 1ull << 62;

__device__ __forceinline__ unsigned long long to_fixed(float val, int e) {
    const double x = ldexp((double)val, -e);
    return (unsigned long long)__double_as_longlong(x + 4503599627370496.0) & ((1ull << 52) - 1);
}

// Add the contribution val of a pixel (skipped unless val > 0, as the
// reference's val > 0 test) to its fixed-point accumulator.  The integer
// addition is exact and order-independent; a +inf value adds 0 and or-s the
// marker in, which only wavefronts holding one execute.
template <class Ptr>
__device__ __forceinline__ void fixed_deposit(Ptr acc, float val, int e) {
    if (!(val > 0.0f)) return;
    atomicAdd(acc, to_fixed(val, e));
    const bool inf = !(val < INFINITY);
    if (__any(inf) && inf) atomicOr(acc, kInfMark);
}

__device__ __forceinline__ float from_fixed(unsigned long long v, int e) {
    if (v & kInfMark) return INFINITY;
    return (float)ldexp((double)(long long)v, e);
}

// One record decoded for the tile with origin (X0, Y0) and last pixel (X1, Y1):
// its footprint box = window n tile n ellipse bounding box, and the chord
// parameters of its rows (see row_span).
struct TileRec {
    Frame f;            // ix, iy absolute
    float e;
    int32_t lx0, ly0;   // box origin in tile coordinates
    int32_t kx0, ky0;   // box origin relative to (ix, iy)
    int32_t bw, area;   // box width and area (0: empty)
    int32_t rows;       // box height
    float A, invA, det, P;  // q = A dx^2 + 2 P dy dx + Q dy^2, det = a00 a11 - a01 a10
    bool full;          // no usable chords: walk whole box rows
};

// The pixels of box row ky (offset from iy) that can have q < 1, as a column
// range [xa, xb] (offsets from ix) within the box; xa > xb when none.  For
// fixed dy, q is a quadratic in dx with minimum (det dy)^2 / A at
// dx* = -P dy / A (Lagrange's identity), so the chord is dx* +- w with
// w = sqrt(A - (det dy)^2) / A.  The chord is widened (1e-5 A under the root,
// 1e-3 px + 1e-4 relative outside) far beyond the float error of q near 1, so
// every pixel the exact evaluation would keep is inside; pixels left out have
// q >= 1 and contribute exactly nothing, and the walk over the chords visits
// the kept pixels in the same order as the walk over the whole box.
__device__ __forceinline__ void row_span(const TileRec& t, int ky, int& xa, int& xb) {
    const int bx1 = t.kx0 + t.bw - 1;
    if (t.full) {
        xa = t.kx0;
        xb = bx1;
        return;
    }
    // explicitly rounded operations: the sort and the walk evaluate this in
    // separate inlined copies and must get the same chord
    const float dy = __fsub_rn((float)ky, t.f.fy);
    const float dd = __fmul_rn(t.det, dy);
    const float rad = fmaf(-dd, dd, t.A);
    const float eps = __fmul_rn(1e-5f, t.A);
    if (!(rad > -eps)) {
        xa = 1;
        xb = 0;
        return;
    }
    const float w = __fmul_rn(__builtin_amdgcn_sqrtf(__fadd_rn(fmaxf(rad, 0.0f), eps)), t.invA);
    const float c = fmaf(__fmul_rn(-t.P, dy), t.invA, t.f.fx);
    const float m = fmaf(1e-4f, __fadd_rn(fabsf(c), w), 1e-3f);
    const float lo = __fsub_rn(__fsub_rn(c, w), m), hi = __fadd_rn(__fadd_rn(c, w), m);
    // clamp in float before converting (c, w may be huge for a thin footprint)
    xa = (int)ceilf(fminf(fmaxf(lo, (float)t.kx0), (float)(bx1 + 1)));
    xb = (int)floorf(fmaxf(fminf(hi, (float)bx1), (float)(t.kx0 - 1)));
}

// Pixels of the record's chords (the walk's length); exact box area when full.
__device__ __forceinline__ int chord_pixels(const TileRec& t) {
    if (t.area == 0) return 0;
    if (t.full) return t.area;
    int n = 0;
    for (int r = 0; r < t.rows; ++r) {
        int xa, xb;
        row_span(t, t.ky0 + r, xa, xb);
        n += max(xb - xa + 1, 0);
    }
    return n;
}

// The walk over a record's chords, one pixel per step: current row r, column
// kx (offset from ix), end of the row's chord xb, LDS index li, row offset dy.
struct ChordWalk {
    int r, kx, xb, li;
    float dy;
    __device__ __forceinline__ float dx(float fx) const { return (float)kx - fx; }
    // move to the first pixel of the first non-empty chord at or after row r
    __device__ __forceinline__ void seek(const TileRec& t, int tshift) {
        int xa = 1;
        xb = 0;
        for (; r < t.rows; ++r) {
            row_span(t, t.ky0 + r, xa, xb);
            if (xa <= xb) break;
        }
        if (r >= t.rows) {  // not reached when the walk's length is the chord count: stay inside the box
            r = t.rows - 1;
            xa = xb = t.kx0;
        }
        kx = xa;
        li = ((t.ly0 + r) << tshift) + t.lx0 + (xa - t.kx0);
        dy = (float)(t.ky0 + r) - t.f.fy;
    }
    __device__ __forceinline__ void start(const TileRec& t, int tshift) {
        r = 0;
        seek(t, tshift);
    }
    // after a pixel: next column, or the next non-empty chord
    __device__ __forceinline__ void step(const TileRec& t, int tshift, bool more) {
        ++kx;
        ++li;
        if (kx > xb && more) {
            ++r;
            seek(t, tshift);
        }
    }
};

// The walk over a record's whole box rows (the default), two pixels of a row
// per step: (kx, ky) and (kx + 1, ky), the second absent (second() false)
// where a row of odd width ends.  A pair shares its row's products a01 dy and
// a11 dy and pays the branch-free row advance once (a divergent advance would
// cost every lane of the wavefront whenever one lane changes row, which is
// most steps).  Per pixel the arithmetic is ChordWalk's (the same dx, dy, u,
// v, q), and the pixels come in the same order.
//
// The pixel offsets (kx, ky) are counted in float: they are small integers,
// exact in float, so (kx - fx) is the same float as ((float)kx - fx) and the
// walk spends no int-to-float conversions per pixel.
struct PairWalk {
    float kx, ky, kx0, xb, xb1;  // xb: last column; a pair at kx >= xb1 = xb - 1 ends its row
    int li, skip;                // LDS index of the pair's first pixel; its step to the next row
    __device__ __forceinline__ void start(const TileRec& t, int tshift) {
        kx0 = (float)t.kx0;
        kx = kx0;
        ky = (float)t.ky0;
        xb = (float)(t.kx0 + t.bw - 1);
        xb1 = xb - 1.0f;
        li = (t.ly0 << tshift) + t.lx0;
        skip = (1 << tshift) - 2 * (((t.bw + 1) >> 1) - 1);
    }
    __device__ __forceinline__ bool second() const { return kx < xb; }
    __device__ __forceinline__ void step() {
        const bool nl = kx >= xb1;
        kx = nl ? kx0 : kx + 2.0f;
        li += nl ? skip : 2;
        ky += nl ? 1.0f : 0.0f;
    }
};

// steps of the pair walk: box rows x pairs per row
__device__ __forceinline__ int walk_pairs(const TileRec& t) { return t.rows * ((t.bw + 1) >> 1); }

// The forward's value at a pixel, with the Silverman kernel's 3/pi folded into
// e3 = (3/pi) E: (1 - q)^2 e3 for q < 1, else 0 (also for a NaN q: fmaxf
// returns its number operand).  As in the reference (photon_differentials.cu:
// 105-106) only val > 0 is splatted, so E <= 0 and NaN energies add nothing.
__device__ __forceinline__ float splat_val(float q, float e3) {
    const float t = fmaxf(1.0f - q, 0.0f);
    return t * t * e3;
}

// The backward's per-pixel footprint weight w = 1 - q inside, 0 outside (and
// for a NaN q), and the moments of one record's walk:  m0 = sum w^2 g,
// m1 = sum w g dx, m2 = sum w g dy, m3 = sum w g dx^2, m4 = sum w g dx dy,
// m5 = sum w g dy^2.
__device__ __forceinline__ float foot_w(float q) { return fmaxf(1.0f - q, 0.0f); }

struct Moments {
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f;
    // a pixel of weight w and weighted gradient wg = w g (0 adds exactly nothing)
    __device__ __forceinline__ void add(float w, float wg, float dx, float dy) {
        m0 = fmaf(w, wg, m0);
        const float px = wg * dx, py = wg * dy;
        m1 += px;
        m2 += py;
        m3 = fmaf(px, dx, m3);
        m4 = fmaf(px, dy, m4);
        m5 = fmaf(py, dy, m5);
    }
};

// Chord parameters (the chord walk only).
__device__ __forceinline__ void chord_setup(TileRec& t, bool box_walk) {
    t.full = true;  // the box walk visits whole box rows: no chord parameters
    t.A = t.invA = t.det = t.P = 0.0f;
    if (!box_walk) {
        const Frame& f = t.f;
        t.A = fmaf(f.a00, f.a00, f.a10 * f.a10);
        t.P = fmaf(f.a00, f.a01, f.a10 * f.a11);
        t.det = det2(f.a00, f.a01, f.a10, f.a11);
        t.invA = 1.0f / t.A;
        t.full = !(t.A > 1e-30f) || !(t.A < 1e30f) || !(fabsf(t.det) < INFINITY) || !(fabsf(t.P) < INFINITY) ||
                 !(t.invA < INFINITY);
    }
}

// One record decoded for the tile with origin (X0, Y0): its frame, energy and
// footprint box (window n tile n ellipse bounding box, made by the scatter).
__device__ __forceinline__ TileRec decode_tile_rec_q(const uint4& q0, const uint4& q1, int X0, int Y0,
                                                     bool box_walk) {
    TileRec t;
    const uint32_t w = q0.w;
    t.lx0 = (int)(q0.x & 63);
    t.ly0 = (int)(q0.y & 63);
    t.bw = (int)(w & 63) + 1;
    t.rows = (int)((w >> 6) & 63) + 1;
    t.area = t.bw * t.rows;
    t.kx0 = (int)((w >> 12) & 255) - 64;
    t.ky0 = (int)((w >> 20) & 255) - 64;
    t.f.ix = X0 + t.lx0 - t.kx0;
    t.f.iy = Y0 + t.ly0 - t.ky0;
    t.f.fx = frac_of(q0.x);
    t.f.fy = frac_of(q0.y);
    t.e = __uint_as_float(q0.z);
    t.f.a00 = __uint_as_float(q1.x);
    t.f.a01 = __uint_as_float(q1.y);
    t.f.a10 = __uint_as_float(q1.z);
    t.f.a11 = __uint_as_float(q1.w);
    chord_setup(t, box_walk);
    return t;
}

__device__ __forceinline__ TileRec decode_tile_rec(const uint4* __restrict__ recs, size_t j, int X0, int Y0,
                                                   bool box_walk) {
    return decode_tile_rec_q(recs[2 * j], recs[2 * j + 1], X0, Y0, box_walk);
}

// Counting sort of a segment's records by their walk length (steps of the
// pair walk: box rows x pixel pairs per row; or the pixels of their chords;
// clamped to 255, the key is exact below that),
// so that the 64 records a wavefront walks side by side, one per lane, have
// nearly the same length.  The box walk's length is in each record's w word:
// that word is all this pass reads.
struct SegSort {
    uint8_t key[kMaxSeg];
    uint16_t order[kMaxSeg];
    uint32_t hist[kSortKeys];
    uint32_t wsum[kSortKeys / 64];
};

__device__ __forceinline__ void sort_segment(SegSort& S, const uint4* __restrict__ recs, const int4& d, int X0, int Y0,
                                             bool box_walk) {
    static_assert(kTileThreads >= kSortKeys && kSortKeys % 64 == 0, "one histogram bin per thread of the scan");
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const bool scan = tid < kSortKeys;  // whole wavefronts
    if (scan) S.hist[tid] = 0;
    __syncthreads();
    // at most kMaxSeg / kTileThreads records per thread, all loaded before any
    // key is computed so that the loads are in flight together
    constexpr int R = kMaxSeg / kTileThreads;
    if (box_walk) {
        const uint32_t* words = reinterpret_cast<const uint32_t*>(recs);
        uint32_t ww[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = tid + r * kTileThreads;
            ww[r] = j < d.z ? words[8 * ((size_t)d.y + j) + 3] : 0u;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = tid + r * kTileThreads;
            if (j < d.z) {
                const uint32_t key = (uint32_t)min(rec_pairs(ww[r]), 255);
                S.key[j] = (uint8_t)key;
                atomicAdd(&S.hist[key], 1u);
            }
        }
    } else {
        uint4 q0[R], q1[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = tid + r * kTileThreads;
            if (j < d.z) {
                q0[r] = recs[2 * ((size_t)d.y + j)];
                q1[r] = recs[2 * ((size_t)d.y + j) + 1];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = tid + r * kTileThreads;
            if (j < d.z) {
                const TileRec t = decode_tile_rec_q(q0[r], q1[r], X0, Y0, false);
                const uint32_t key = (uint32_t)min(chord_pixels(t), 255);
                S.key[j] = (uint8_t)key;
                atomicAdd(&S.hist[key], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t v = 0, incl = 0;
    if (scan) {
        v = S.hist[tid];
        incl = (uint32_t)wave_inclusive_sum((int64_t)v);
        if (lane == 63) S.wsum[w] = incl;
    }
    __syncthreads();
    if (scan) {
        uint32_t excl = incl - v;
        for (int k = 0; k < w; ++k) excl += S.wsum[k];
        S.hist[tid] = excl;
    }
    __syncthreads();
    for (int j = tid; j < d.z; j += kTileThreads) {
        const uint32_t pos = atomicAdd(&S.hist[S.key[j]], 1u);
        S.order[pos] = (uint16_t)j;
    }
    __syncthreads();
}

// The walk order a forward keeps for the tile adjoint: entries record |
// key << 16, grouped by storage part (records j of part j / kStageRecs, in
// order of the part), each group in walk order (a stable partition of the
// sorted positions), so that the adjoint walks one part at a time and can
// stage that part's partials in LDS.  Every part but the last is full, so part
// p starts at p kStageRecs.  Each thread partitions kMaxSeg / kTileThreads
// consecutive sorted positions; a workgroup scan of the threads' per-part
// counts (16-bit fields of one 64-bit sum) places them.
__device__ __forceinline__ void write_walk_order(const SegSort& S, const int4& d, uint32_t* __restrict__ order_out) {
    constexpr int R = kMaxSeg / kTileThreads;
    constexpr int kShift = ilog2_c(kStageRecs);
    __shared__ int64_t wsum[kTileThreads / 64];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    uint32_t ent[R];
    int64_t cnt = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pos = tid * R + r;
        ent[r] = 0u;
        if (pos < d.z) {
            const uint32_t j = S.order[pos];
            ent[r] = j | ((uint32_t)S.key[j] << 16);
            cnt += (int64_t)1 << (16 * (j >> kShift));
        }
    }
    const int64_t incl = wave_inclusive_sum(cnt);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int64_t rank = incl - cnt;  // per part: its entries at sorted positions before this thread's
    for (int k = 0; k < w; ++k) rank += wsum[k];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pos = tid * R + r;
        if (pos < d.z) {
            const int part = (int)((ent[r] & 0xffffu) >> kShift);
            const int dst = part * kStageRecs + (int)((rank >> (16 * part)) & 0xffff);
            order_out[(size_t)d.y + dst] = ent[r];
            rank += (int64_t)1 << (16 * part);
        }
    }
}

// Wavefront work queue over a segment's sorted positions: chunks of 64
// consecutive positions, handed out largest walks first (the sort is
// ascending), so that the workgroup's wavefronts finish together.  A
// static round-robin split gives the last wavefront the largest records of
// every round and leaves the others waiting at the flush barrier.
// Returns the chunk index, or -1 when the queue is empty.
__device__ __forceinline__ int next_chunk(int* counter, int nchunks) {
    int c = 0;
    if ((threadIdx.x & 63) == 0) c = atomicAdd(counter, 1);
    c = __shfl(c, 0);
    return c < nchunks ? nchunks - 1 - c : -1;
}

// One lane's record of a queue chunk: chunk index, walk length key (0:
// nothing to walk, including lanes past the segment's end) and the record's
// two halves.
struct ChunkRecs {
    int ch, key;
    uint4 q0, q1;
};

__device__ __forceinline__ ChunkRecs load_chunk(const SegSort& S, const uint4* __restrict__ recs, const int4& d,
                                                int ch) {
    ChunkRecs c;
    c.ch = ch;
    c.key = 0;
    c.q0 = c.q1 = make_uint4(0u, 0u, 0u, 0u);
    if (ch < 0) return c;
    const int pos = (ch << 6) + (threadIdx.x & 63);
    if (pos >= d.z) return c;
    const int j = S.order[pos];
    c.key = (int)S.key[j];
    if (c.key == 0) return c;
    c.q0 = recs[2 * ((size_t)d.y + j)];
    c.q1 = recs[2 * ((size_t)d.y + j) + 1];
    return c;
}

constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, each with its own L2

// Segment geometry: channel and tile of the segment's bin.
struct SegTile {
    int c, X0, Y0, X1, Y1;
};

__device__ __forceinline__ SegTile seg_tile(const Geom& g, const ImageArgs& I, int bin) {
    SegTile st;
    const int tpc = g.tiles_x * g.tiles_y;
    st.c = bin / tpc;
    const int tt = bin - st.c * tpc;
    const int ty = tt / g.tiles_x, tx = tt - ty * g.tiles_x;
    st.X0 = tx << g.tshift;
    st.Y0 = ty << g.tshift;
    st.X1 = min(st.X0 + g.T, I.W) - 1;
    st.Y1 = min(st.Y0 + g.T, I.H) - 1;
    return st;
}

// Photon-order forward of every photon (flattened walk, global atomics or
// global int64 atomics), grid-stride over 64-photon batches, one per wavefront
// of the kTileThreads-thread workgroup.
template <bool DET>
__device__ __forceinline__ void flat_splat_all(const PhotonArgs& P, const ImageArgs& I, int fe, WaveRecs* recs,
                                               float* __restrict__ out, unsigned long long* __restrict__ acc64) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveRecs& R = recs[w];
    R.mark[lane] = -1;
    const int64_t n = live_photons(P);
    const int64_t nbatch = (n + 63) >> 6;
    for (int64_t b = (int64_t)blockIdx.x * (kTileThreads / 64) + w; b < nbatch;
         b += (int64_t)gridDim.x * (kTileThreads / 64)) {
        const int64_t i = (b << 6) + lane;
        FlatRec r;
        setup_flat_rec<kFwdUnit>(P, I, i, /*forward=*/true, r, n);
        const int64_t total = publish_recs(R, r, lane);
        walk_units<kFwdUnit>(R, r, lane, total, [&](const FlatItem& it) {
            if constexpr (DET) {
                const UnitPixels<kFwdUnit> px(it, I.W);
#pragma unroll
                for (int k = 0; k < kFwdUnit; ++k) {
                    if (k < it.n) {
                        float u, v;
                        const float val = silverman(eval_q(it.f, px.dx[k], px.dy[k], u, v)) * it.e;
                        fixed_deposit(&acc64[px.pix[k]], val, fe);
                    }
                }
            } else {
                flat_forward_unit<kFwdUnit>(it, I.W, out);
            }
        });
    }
}

// ---------------------------------------------------------------------------
// 4. tile (forward): one workgroup per segment; each lane splats whole
//    records (sorted by walk length) into the LDS tile, which is then flushed.  On
//    overflow the workgroups run the photon-order walk instead.
// ---------------------------------------------------------------------------

// Eight wavefronts per SIMD (four workgroups per CU, as the LDS allows): the
// deterministic instance otherwise takes 104 SGPRs, which leaves room for
// seven, i.e. three workgroups per CU; with the bound it spills 14 SGPRs into
// VGPR lanes and its kernel time on the benchmark splat went 85.1 -> 78.0 us.
// The atomic instance fits eight as it is (100 SGPRs); bounded too, its code
// fetches 56 MB more per launch (from the last-level cache, it seems) but the
// forward measured ~3 us shorter than without the bound.
template <bool DET>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8))) void bin_tile_kernel(
    Geom g, PhotonArgs P, ImageArgs I, const int4* __restrict__ segs, const uint4* __restrict__ recs,
    const Ctrl* ctrl, float* __restrict__ out, unsigned long long* __restrict__ acc64,
    uint32_t* __restrict__ order_out) {
    // LDS accumulators: int64 fixed point (deterministic) or double.  Both
    // have fast LDS atomics on gfx950 (ds_add_u64, ds_add_f64), whereas the
    // single-precision ds_add_f32 serialises its lanes (about 3 cycles per
    // active lane per CU), which made it the tile kernel's bottleneck.
    using Acc = typename std::conditional<DET, unsigned long long, double>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ SegSort S;
    Acc* tile = reinterpret_cast<Acc*>(smem);
    if (ctrl->overflow) {
        // the dynamic LDS (sized for both) holds the walk's per-wave buffers
        flat_splat_all<DET>(P, I, DET ? ctrl->fx_exp : 0, reinterpret_cast<WaveRecs*>(smem), out, acc64);
        return;
    }
    const uint32_t s = blockIdx.x;
    if (s >= ctrl->nseg) return;
    const int4 d = segs[s];
    const int fe = DET ? ctrl->fx_exp : 0;
    const int T = g.T, TT = T * T;
    __shared__ int wq;  // next_chunk's counter
    if (threadIdx.x == 0) wq = 0;
    for (int k = threadIdx.x; k < TT; k += kTileThreads) tile[k] = Acc(0);
    const SegTile st = seg_tile(g, I, d.x);
    sort_segment(S, recs, d, st.X0, st.Y0, g.box_walk);
    if (order_out) write_walk_order(S, d, order_out);  // for the backward
    const int nchunks = (d.z + 63) >> 6;
    // (loading the next chunk's records while the current one walks measured
    // slower: 81 against 74 us, the registers they hold cost occupancy)
    for (int ch; (ch = next_chunk(&wq, nchunks)) >= 0;) {
        const ChunkRecs cur = load_chunk(S, recs, d, ch);
        const int key = cur.key;
        if (key == 0) continue;
        const TileRec t = decode_tile_rec_q(cur.q0, cur.q1, st.X0, st.Y0, g.box_walk);
        const int n = key < 255 ? key : (g.box_walk ? walk_pairs(t) : chord_pixels(t));
        const float e3 = kThreeOverPi * t.e;
        auto deposit = [&](int li, float val) {
            if constexpr (DET)
                fixed_deposit(&tile[li], val, fe);
            else if (val > 0.0f)
                atomicAdd(&tile[li], (double)val);
        };
        if (g.box_walk) {
            PairWalk w;
            w.start(t, g.tshift);
            for (int p = 0; p < n; ++p) {
                const float dy = w.ky - t.f.fy;
                const float ru = t.f.a01 * dy, rv = t.f.a11 * dy;
                const float dx0 = w.kx - t.f.fx, dx1 = (w.kx + 1.0f) - t.f.fx;
                const float u0 = fmaf(t.f.a00, dx0, ru), v0 = fmaf(t.f.a10, dx0, rv);
                const float u1 = fmaf(t.f.a00, dx1, ru), v1 = fmaf(t.f.a10, dx1, rv);
                const float val0 = splat_val(fmaf(u0, u0, v0 * v0), e3);
                const float val1 = w.second() ? splat_val(fmaf(u1, u1, v1 * v1), e3) : 0.0f;
                deposit(w.li, val0);
                deposit(w.li + 1, val1);
                w.step();
            }
        } else {
            ChordWalk w;
            if (n > 0) w.start(t, g.tshift);
            for (int p = 0; p < n; ++p) {
                float u, v;
                deposit(w.li, splat_val(eval_q(t.f, w.dx(t.f.fx), w.dy, u, v), e3));
                w.step(t, g.tshift, p + 1 < n);
            }
        }
    }
    __syncthreads();
    const bool split = (d.w & 1) != 0;
    for (int k = threadIdx.x; k < TT; k += kTileThreads) {
        const int px = st.X0 + (k & (T - 1)), py = st.Y0 + (k >> g.tshift);
        if (px >= I.W || py >= I.H) continue;
        const int64_t o = ((int64_t)st.c * I.H + py) * I.W + px;
        const Acc v = tile[k];
        if constexpr (DET) {
            if (!split) {
                out[o] = from_fixed(v, fe);
            } else if (v != 0) {
                // integer addition of the finite part (order-independent); the
                // +inf marker is or-ed in, so that two segments' markers do
                // not carry out of bit 62
                if (v & kInfMark) atomicOr(&acc64[o], kInfMark);
                const unsigned long long fin = v & (kInfMark - 1);
                if (fin) atomicAdd(&acc64[o], fin);
            }
        } else {
            if (!split)
                out[o] = (float)v;
            else if (v != 0.0)
                unsafeAtomicAdd(&out[o], (float)v);
        }
    }
}

// ---------------------------------------------------------------------------
// 5. convert split tiles (deterministic mode), or the whole image after the
//    photon-order path
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTileThreads) void bin_convert_split_kernel(Geom g, ImageArgs I,
                                                                         const int4* __restrict__ segs,
                                                                         const Ctrl* ctrl,
                                                                         const unsigned long long* __restrict__ acc64,
                                                                         float* __restrict__ out) {
    if (ctrl->overflow) {
        const int fe = ctrl->fx_exp;
        const int64_t n = (int64_t)I.C * I.H * I.W;
        const int64_t stride = (int64_t)gridDim.x * kTileThreads;
        for (int64_t k = (int64_t)blockIdx.x * kTileThreads + threadIdx.x; k < n; k += stride)
            out[k] = from_fixed(acc64[k], fe);
        return;
    }
    const uint32_t s = blockIdx.x;
    if (s >= ctrl->nseg) return;
    const int4 d = segs[s];
    if ((d.w & 3) != 3) return;
    const int tpc = g.tiles_x * g.tiles_y;
    const int c = d.x / tpc, t = d.x - c * tpc;
    const int ty = t / g.tiles_x, tx = t - ty * g.tiles_x;
    const int X0 = tx << g.tshift, Y0 = ty << g.tshift;
    const int fe = ctrl->fx_exp;
    for (int k = threadIdx.x; k < g.T * g.T; k += kTileThreads) {
        const int px = X0 + (k & (g.T - 1)), py = Y0 + (k >> g.tshift);
        if (px < I.W && py < I.H) {
            const int64_t o = ((int64_t)c * I.H + py) * I.W + px;
            out[o] = from_fixed(acc64[o], fe);
        }
    }
}

// ---------------------------------------------------------------------------
// One record's partial gradient: the reference's seven sums over the record's
// pixels (photon_differentials.cu:195-204), in the reference's units, from
// the moments of its walk of n steps against the gradient tile gt (last:
// its last index).  With s = dk / r = c w g, c = -(12/pi) E, and
// (u, v) = A (dx, dy), the sums are linear in the moments (Moments).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void record_partial(const TileRec& t, int n, const float* gt, int last, int tshift,
                                               bool box_walk, float hw, float hh, float iw, float ih, float4& p0,
                                               float4& p1) {
    Moments m;
    if (box_walk) {
        PairWalk cw;
        cw.start(t, tshift);
        for (int p = 0; p < n; ++p) {
            const float dy = cw.ky - t.f.fy;
            const float ru = t.f.a01 * dy, rv = t.f.a11 * dy;
            const float dx0 = cw.kx - t.f.fx, dx1 = (cw.kx + 1.0f) - t.f.fx;
            const float u0 = fmaf(t.f.a00, dx0, ru), v0 = fmaf(t.f.a10, dx0, rv);
            const float u1 = fmaf(t.f.a00, dx1, ru), v1 = fmaf(t.f.a10, dx1, rv);
            const float w0 = foot_w(fmaf(u0, u0, v0 * v0)), w1 = foot_w(fmaf(u1, u1, v1 * v1));
            // the second pixel of a pair may lie past the row (past the tile
            // at its last pixel): masked, not multiplied
            const float g0 = gt[cw.li], g1 = gt[min(cw.li + 1, last)];
            m.add(w0, w0 * g0, dx0, dy);
            m.add(w1, cw.second() ? w1 * g1 : 0.0f, dx1, dy);
            cw.step();
        }
    } else {
        ChordWalk cw;
        if (n > 0) cw.start(t, tshift);
        for (int p = 0; p < n; ++p) {
            const float dx = cw.dx(t.f.fx);
            float u, v;
            const float w = foot_w(eval_q(t.f, dx, cw.dy, u, v));
            m.add(w, w * gt[cw.li], dx, cw.dy);
            cw.step(t, tshift, p + 1 < n);
        }
    }
    const float m0 = m.m0, m1 = m.m1, m2 = m.m2, m3 = m.m3, m4 = m.m4, m5 = m.m5;
    const Frame& f = t.f;
    const float c = -kTwelveOverPi * t.e;
    const float su = c * fmaf(f.a00, m1, f.a01 * m2);  // sum s u
    const float sv = c * fmaf(f.a10, m1, f.a11 * m2);  // sum s v
    const float sudx = c * fmaf(f.a00, m3, f.a01 * m4), sudy = c * fmaf(f.a00, m4, f.a01 * m5);
    const float svdx = c * fmaf(f.a10, m3, f.a11 * m4), svdy = c * fmaf(f.a10, m4, f.a11 * m5);
    const float gx0 = -fmaf(f.a00, su, f.a10 * sv), gx1 = -fmaf(f.a01, su, f.a11 * sv);
    p0 = make_float4(kThreeOverPi * m0, gx0 * hw, gx1 * hh, sudx * iw);
    p1 = make_float4(sudy * ih, svdx * iw, svdy * ih, 0.0f);
}

// ---------------------------------------------------------------------------
// backward 1: tile adjoint.  The segment's gradient tile is staged in LDS; each
// lane walks whole records (sorted by walk length), sums their seven terms in
// registers and stores them, in the reference's units, at the record's slot.
// ---------------------------------------------------------------------------
// SORTED: the forward of the same binning left each segment's walk order (and
// every record's walk length) in order_in, grouped by storage part
// (write_walk_order), so the segment is not sorted again, and the partials of
// one part at a time are staged in LDS and then stored coalesced: stored
// straight from the walk, at positions in walk order, they were scattered
// 32-byte writes whose lines the L2 evicted half-written (backward pass
// 165.5 -> 160.5 us on the benchmark splat, four alternating runs).
template <bool SORTED>
__global__ __launch_bounds__(kTileThreads) void bin_tile_bwd_kernel(Geom g, ImageArgs I, const int4* __restrict__ segs,
                                                                    const uint4* __restrict__ recs, const Ctrl* ctrl,
                                                                    const float* __restrict__ grad,
                                                                    float4* __restrict__ partials,
                                                                    const uint32_t* __restrict__ order_in) {
    extern __shared__ __align__(16) float gt[];  // T * T upstream gradient tile
    const uint32_t s = blockIdx.x;
    if (s >= ctrl->nseg) return;
    const int4 d = segs[s];
    const int T = g.T, TT = T * T;
    const SegTile st = seg_tile(g, I, d.x);
    for (int k = threadIdx.x; k < TT; k += kTileThreads) {
        const int px = st.X0 + (k & (T - 1)), py = st.Y0 + (k >> g.tshift);
        gt[k] = (px < I.W && py < I.H) ? grad[((int64_t)st.c * I.H + py) * I.W + px] : 0.0f;
    }
    constexpr int kWaves = kTileThreads / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float hw = 0.5f * I.fw, hh = 0.5f * I.fh, iw = 2.0f / I.fw, ih = 2.0f / I.fh;
    __shared__ int wq;  // next_chunk's counter
    if constexpr (SORTED) {
        __shared__ float4 stg[2 * kStageRecs];  // one storage part's partials
        for (int part = 0; part < kStageParts; ++part) {
            // order_in positions of this part: [base, base + cnt), records base + [0, cnt)
            const int base = part * kStageRecs, cnt = min(d.z - base, kStageRecs);
            if (cnt <= 0) break;
            const int nch = (cnt + 63) >> 6;
            // Each wavefront's first chunk is taken statically (the kWaves
            // largest walks, one each), the rest from the queue, whose counter
            // starts past them.
            if (threadIdx.x == 0) wq = kWaves;
            // the gradient tile is loaded (first part); the staged partials of
            // the previous part are stored (later parts)
            __syncthreads();
            for (int ch = wv < nch ? nch - 1 - wv : -1; ch >= 0; ch = next_chunk(&wq, nch)) {
                const int pos = (ch << 6) + lane;
                if (pos >= cnt) continue;
                const uint32_t w = order_in[(size_t)d.y + base + pos];
                const int j = (int)(w & 0xffffu), key = (int)(w >> 16);
                const TileRec t = decode_tile_rec_q(recs[2 * ((size_t)d.y + j)], recs[2 * ((size_t)d.y + j) + 1],
                                                    st.X0, st.Y0, g.box_walk);
                const int n = key < 255 ? key : (g.box_walk ? walk_pairs(t) : chord_pixels(t));
                const int k = j - base;  // in [0, cnt): each once
                record_partial(t, n, gt, TT - 1, g.tshift, g.box_walk, hw, hh, iw, ih, stg[2 * k], stg[2 * k + 1]);
            }
            __syncthreads();
            float4* dst = partials + 2 * ((size_t)d.y + base);
            for (int k = threadIdx.x; k < 2 * cnt; k += kTileThreads) dst[k] = stg[k];
        }
    } else {
        const int nchunks = (d.z + 63) >> 6;
        // Each wavefront's first chunk is taken statically (the kWaves largest
        // walks, one each), the rest from the queue, whose counter starts past them.
        if (threadIdx.x == 0) wq = kWaves;
        const int first = wv < nchunks ? nchunks - 1 - wv : -1;
        __shared__ SegSort S;
        sort_segment(S, recs, d, st.X0, st.Y0, g.box_walk);
        // one lane's record of a chunk: record index (-1 for lanes past the
        // segment's end), walk length key and the record's two halves
        auto load = [&](int ch, int& j, int& key, uint4& q0, uint4& q1) {
            key = 0;
            j = -1;
            const int pos = (ch << 6) + lane;
            if (ch < 0 || pos >= d.z) return;
            j = S.order[pos];
            key = (int)S.key[j];
            q0 = recs[2 * ((size_t)d.y + j)];
            q1 = recs[2 * ((size_t)d.y + j) + 1];
        };
        int ch = first, j = 0, key = 0;
        uint4 q0 = make_uint4(0u, 0u, 0u, 0u), q1 = q0;
        load(ch, j, key, q0, q1);
        // partials staged per wavefront and written by lane pairs (one 32-byte
        // request per record, as in the scatter), at the records' positions
        __shared__ float4 pst[kTileThreads / 64][2][64];
        __shared__ uint32_t pj[kTileThreads / 64][64];
        for (; ch >= 0; ch = next_chunk(&wq, nchunks), load(ch, j, key, q0, q1)) {
            const bool act = j >= 0;
            if (act) {
                const TileRec t = decode_tile_rec_q(q0, q1, st.X0, st.Y0, g.box_walk);
                const int n = key < 255 ? key : (g.box_walk ? walk_pairs(t) : chord_pixels(t));
                record_partial(t, n, gt, TT - 1, g.tshift, g.box_walk, hw, hh, iw, ih, pst[wv][0][lane],
                               pst[wv][1][lane]);
                pj[wv][lane] = (uint32_t)j;
            }
            const uint64_t mask = __ballot(act);
            wave_lds_fence();
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int src = (lane >> 1) + 32 * p, h = lane & 1;
                if ((mask >> src) & 1ull) partials[2 * ((size_t)d.y + pj[wv][src]) + h] = pst[wv][h][src];
            }
            wave_lds_fence();
        }
    }
}

// ---------------------------------------------------------------------------
// backward 2: finalize in photon order, one thread per photon -- sum each
// photon's record partials (in tile order) and store coalesced.  Special
// photons (non-finite inputs, more than kMaxRecsPerPhoton records), and every
// photon when the records overflowed the workspace, are evaluated here by a
// direct serial walk instead (NaN/inf propagate as in the reference).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bin_finalize_kernel(const float* __restrict__ grad, PhotonArgs P, ImageArgs I,
                                                           const Ctrl* ctrl, const uint8_t* __restrict__ info,
                                                           const uint32_t* __restrict__ slots,
                                                           const float4* __restrict__ partials, GradArgs G,
                                                           int32_t nblk, int64_t rounds) {
    // XCD-aware photon order.  Workgroups are dispatched round-robin over the
    // 8 XCDs (workgroup b on XCD b % 8); the half chunks (256 photons) of one
    // binning workgroup go to workgroups of one XCD, at consecutive positions
    // in that XCD's sequence.  A binning workgroup's records lie in short
    // runs (one per bin), so its photons' 32-byte partials share cache lines,
    // which are then fetched once into that XCD's L2 instead of once per XCD.
    static_assert(kChunk == 512, "two finalize workgroups per chunk");
    const int64_t N = P.N;
    const int64_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
    const int64_t blk = x + kXcds * (k / (2 * rounds)), h = k % (2 * rounds);  // binning workgroup, half chunk
    const int64_t i = ((h >> 1) * nblk + blk) * kChunk + (h & 1) * 256 + threadIdx.x;
    // photons beyond a batch's live count are not written
    if (blk >= nblk || i >= live_photons(P)) return;
    // on overflow the scatter wrote no records, slots or classes
    const uint32_t cls = ctrl->overflow ? (uint32_t)kSpecial : (uint32_t)stream_load(&info[i]);
    if (cls == kSpecial) {
        float a[7];
        photon_grad_direct(grad, P, I, i, /*clip_box=*/true, a);
        store_grads(G, I, N, i, a);
        return;
    }
    // all slot loads first, then all partial loads: up to 8 gathers in flight
    uint32_t j[kMaxRecsPerPhoton];
#pragma unroll
    for (int k = 0; k < kMaxRecsPerPhoton; ++k) j[k] = k < (int)cls ? stream_load(&slots[(size_t)k * N + i]) : 0u;
    float4 p[2 * kMaxRecsPerPhoton];
#pragma unroll
    for (int k = 0; k < kMaxRecsPerPhoton; ++k) {
        if (k < (int)cls) {
            p[2 * k] = partials[2 * (size_t)j[k]];
            p[2 * k + 1] = partials[2 * (size_t)j[k] + 1];
        } else {
            p[2 * k] = p[2 * k + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    float4 s0 = p[0], s1 = p[1];
#pragma unroll
    for (int k = 1; k < kMaxRecsPerPhoton; ++k) {
        s0.x += p[2 * k].x;
        s0.y += p[2 * k].y;
        s0.z += p[2 * k].z;
        s0.w += p[2 * k].w;
        s1.x += p[2 * k + 1].x;
        s1.y += p[2 * k + 1].y;
        s1.z += p[2 * k + 1].z;
    }
    stream_store(s0.x, &G.gE[i]);
    stream_store(s0.y, &G.gx[i]);
    stream_store(s0.z, &G.gx[N + i]);
    stream_store(s0.w, &G.gM[i]);
    stream_store(s1.x, &G.gM[N + i]);
    stream_store(0.0f, &G.gM[2 * N + i]);
    stream_store(s1.y, &G.gM[3 * N + i]);
    stream_store(s1.z, &G.gM[4 * N + i]);
    stream_store(0.0f, &G.gM[5 * N + i]);
}

// ---------------------------------------------------------------------------
// photon-order deterministic forward (problems the binned path cannot take)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void guarded_fill64_kernel(const Ctrl* ctrl, unsigned long long* __restrict__ p,
                                                             int64_t n) {
    if (!ctrl->overflow) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

__global__ __launch_bounds__(256) void guarded_convert_kernel(const Ctrl* ctrl, const unsigned long long* __restrict__ p,
                                                              float* __restrict__ out, int64_t n) {
    if (!ctrl->overflow) return;
    const int fe = ctrl->fx_exp;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = from_fixed(p[i], fe);
}

template <bool DET>
__global__ __launch_bounds__(256) void guarded_flat_kernel(PhotonArgs P, ImageArgs I, const Ctrl* ctrl,
                                                           float* __restrict__ out,
                                                           unsigned long long* __restrict__ acc64) {
    __shared__ WaveRecs recs[4];
    if (!ctrl->overflow) return;
    flat_splat_all<DET>(P, I, ctrl->fx_exp, recs, out, acc64);
}

// maximum finite positive energy of the photons that can contribute
__global__ __launch_bounds__(256) void emax_kernel(PhotonArgs P, ImageArgs I, Ctrl* ctrl) {
    __shared__ float wmax[4];
    float m = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.N; i += stride) {
        const int64_t c = P.cp[i];
        const float e = P.E[i];
        if (c >= 0 && c < I.C && e > 0.0f && isfinite(e)) m = fmaxf(m, e);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&ctrl->max_e_bits, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]),
                                                                               fmaxf(wmax[2], wmax[3]))));
}

// fixed-point quantum for the photon-order deterministic path (which the
// guarded kernels run: overflow = 1)
__global__ void fixed_point_kernel(Ctrl* ctrl, int64_t N) {
    const float m = __uint_as_float(ctrl->max_e_bits);
    ctrl->fx_exp = fixed_exponent(N, m);
    ctrl->overflow = 1;
}

// Deterministic forward on images where the reference's window ellipse test
// is active (photon_differentials.cu:93): the plain schedule, int64 atomics.
__global__ __launch_bounds__(256) void fwd_plain_fixed_kernel(PhotonArgs P, ImageArgs I, const Ctrl* ctrl,
                                                              unsigned long long* __restrict__ acc64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    const PlainSetup s = plain_setup(P, I, i);
    if (!s.ok) return;
    const int fe = ctrl->fx_exp;
    unsigned long long* plane = acc64 + (int64_t)s.c * I.H * I.W;
    const Frame& f = s.f;
    for (int64_t py = s.y0; py <= s.y1; ++py) {
        const float dy = (float)((int32_t)py - f.iy) - f.fy;
        for (int64_t px = s.x0; px <= s.x1; ++px) {
            if (!I.ell_trivial && !window_ellipse_ok((int32_t)(px - s.win.pxc), (int32_t)(py - s.win.pyc), I.W, I.H))
                continue;
            const float dx = (float)((int32_t)px - f.ix) - f.fx;
            float u, v;
            const float val = silverman(eval_q(f, dx, dy, u, v)) * s.e;
            fixed_deposit(&plane[py * I.W + px], val, fe);
        }
    }
}

__global__ void ctrl_init_kernel(Ctrl* ctrl) {
    ctrl->nseg = 0;
    ctrl->overflow = 0;
    ctrl->max_e_bits = 0;
    ctrl->fx_exp = 0;
    ctrl->total = 0;
}

}  // namespace

const char* binned_last_error() { return g_err; }

namespace {

struct Buffers {
    uint32_t *hist, *totals, *binoff, *slots;
    float* emax;
    int4* segs;
    Ctrl* ctrl;
    uint4* recs;
    uint8_t* info;
    float4* partials;
    uint32_t* order;
    unsigned long long* acc64;
};

Buffers buffers(void* ws, const Layout& L) {
    unsigned char* b = (unsigned char*)ws;
    Buffers B;
    B.hist = (uint32_t*)(b + L.hist);
    B.totals = (uint32_t*)(b + L.totals);
    B.binoff = (uint32_t*)(b + L.binoff);
    B.emax = (float*)(b + L.emax);
    B.segs = (int4*)(b + L.segs);
    B.ctrl = (Ctrl*)(b + L.ctrl);
    B.recs = (uint4*)(b + L.recs);
    B.slots = (uint32_t*)(b + L.slots);
    B.info = (uint8_t*)(b + L.info);
    B.partials = (float4*)(b + L.partials);
    B.order = (uint32_t*)(b + L.order);
    B.acc64 = (unsigned long long*)(b + L.acc64);
    return B;
}

// Can this problem take the binned path?  (window ellipse test inactive, cap
// <= 63, bins fit the LDS histogram, 32-bit record offsets)
bool binnable(const ImageArgs& I, int64_t N, int32_t mode, const pds_tuning* t, Geom& g) {
    return I.ell_trivial && make_geom(N, I.C, I.H, I.W, I.cap, mode, t, g);
}

// init -> count -> plan -> scatter
void launch_build(const PhotonArgs& P, const ImageArgs& I, const Geom& g, const Buffers& B, bool keep, float* out,
                  hipStream_t s) {
    const size_t hist_lds = sizeof(uint32_t) * (size_t)g.nbins;
    hipLaunchKernelGGL(bin_init_kernel, dim3(1), dim3(1024), 0, s, g, B.totals, B.ctrl);
    hipLaunchKernelGGL(bin_count_kernel, dim3(g.G), dim3(kCountThreads), hist_lds, s, P, I, g, B.hist, B.totals,
                       B.emax);
    if (!g.fused_plan)
        hipLaunchKernelGGL(bin_plan_kernel, dim3(1), dim3(1024), 0, s, g, P.N, (const uint32_t*)B.totals,
                           (const float*)B.emax, B.binoff, B.segs, B.ctrl);
    hipLaunchKernelGGL(keep ? bin_scatter_kernel<true> : bin_scatter_kernel<false>, dim3(g.G), dim3(kScatterThreads),
                       hist_lds, s, P, I, g, (const uint32_t*)B.hist, (const uint32_t*)B.binoff,
                       (const uint32_t*)B.totals, (const float*)B.emax, B.segs, B.recs,
                       keep ? B.slots : (uint32_t*)nullptr, keep ? B.info : (uint8_t*)nullptr, B.ctrl, out, B.acc64);
}

// tiles (-> convert); the same launches run the photon-order path on overflow
void launch_forward_tiles(const PhotonArgs& P, const ImageArgs& I, const Geom& g, const Buffers& B, float* out,
                          bool keep, hipStream_t s) {
    const unsigned nsegs = (unsigned)g.max_segments;
    // the tile accumulator, or (photon-order path) the wavefronts' walk buffers
    const size_t tile_lds = std::max((size_t)g.T * g.T * 8, sizeof(WaveRecs) * (kTileThreads / 64));
    if (g.det) {
        hipLaunchKernelGGL(bin_tile_kernel<true>, dim3(nsegs), dim3(kTileThreads), tile_lds, s, g, P, I,
                           (const int4*)B.segs, (const uint4*)B.recs, (const Ctrl*)B.ctrl, out, B.acc64,
                           keep ? B.order : (uint32_t*)nullptr);
        hipLaunchKernelGGL(bin_convert_split_kernel, dim3(nsegs), dim3(kTileThreads), 0, s, g, I,
                           (const int4*)B.segs, (const Ctrl*)B.ctrl, (const unsigned long long*)B.acc64, out);
    } else {
        hipLaunchKernelGGL(bin_tile_kernel<false>, dim3(nsegs), dim3(kTileThreads), tile_lds, s, g, P, I,
                           (const int4*)B.segs, (const uint4*)B.recs, (const Ctrl*)B.ctrl, out, B.acc64,
                           keep ? B.order : (uint32_t*)nullptr);
    }
}

// tile adjoint -> finalize -> direct walk of special photons (of every photon on overflow)
void launch_backward_tiles(const float* grad, const PhotonArgs& P, const ImageArgs& I, const Geom& g,
                           const Buffers& B, const GradArgs& G, bool sorted, hipStream_t s) {
    const unsigned nsegs = (unsigned)g.max_segments;
    const size_t tile_lds = (size_t)g.T * g.T * 4;  // the gradient tile
    if (sorted)
        hipLaunchKernelGGL(bin_tile_bwd_kernel<true>, dim3(nsegs), dim3(kTileThreads), tile_lds, s, g, I,
                           (const int4*)B.segs, (const uint4*)B.recs, (const Ctrl*)B.ctrl, grad, B.partials,
                           (const uint32_t*)B.order);
    else
        hipLaunchKernelGGL(bin_tile_bwd_kernel<false>, dim3(nsegs), dim3(kTileThreads), tile_lds, s, g, I,
                           (const int4*)B.segs, (const uint4*)B.recs, (const Ctrl*)B.ctrl, grad, B.partials,
                           (const uint32_t*)nullptr);
    // every (binning workgroup, half chunk), binning workgroups rounded up to a multiple of kXcds
    const int64_t blocks = (((int64_t)g.G + kXcds - 1) / kXcds) * kXcds * 2 * g.rounds;
    hipLaunchKernelGGL(bin_finalize_kernel, dim3((unsigned)blocks), dim3(256), 0, s, grad, P, I, (const Ctrl*)B.ctrl,
                       (const uint8_t*)B.info, (const uint32_t*)B.slots, (const float4*)B.partials, G, g.G,
                       g.rounds);
}

// Photon-order forward for problems the binned path cannot take.
int forward_unbinned(const PhotonArgs& P, const ImageArgs& I, bool det, float* out, void* ws, size_t ws_bytes,
                     hipStream_t s) {
    const int64_t numel = (int64_t)I.C * I.H * I.W;
    if (!det) {
        if (launch_fill_zero(out, numel, s) != hipSuccess) return PDS_ERR_HIP;
        if (!I.ell_trivial) return launch_forward_plain(P, I, out, s) == hipSuccess ? PDS_OK : PDS_ERR_HIP;
        return launch_forward_flat(P, I, out, s) == hipSuccess ? PDS_OK : PDS_ERR_HIP;
    }
    // deterministic: int64 fixed point over the whole image
    if (ws == nullptr || ws_bytes < align256(sizeof(Ctrl)) + sizeof(uint64_t) * (size_t)numel) {
        g_err = "workspace too small";
        return PDS_ERR_WORKSPACE;
    }
    unsigned char* base = (unsigned char*)ws;
    Ctrl* ctrl = (Ctrl*)base;
    unsigned long long* acc = (unsigned long long*)(base + align256(sizeof(Ctrl)));
    hipLaunchKernelGGL(ctrl_init_kernel, dim3(1), dim3(1), 0, s, ctrl);
    hipLaunchKernelGGL(emax_kernel, dim3(1024), dim3(256), 0, s, P, I, ctrl);
    hipLaunchKernelGGL(fixed_point_kernel, dim3(1), dim3(1), 0, s, ctrl, P.N);
    hipLaunchKernelGGL(guarded_fill64_kernel, dim3(1024), dim3(256), 0, s, (const Ctrl*)ctrl, acc, numel);
    if (I.ell_trivial)
        hipLaunchKernelGGL(guarded_flat_kernel<true>, dim3(1024), dim3(256), 0, s, P, I, (const Ctrl*)ctrl, out, acc);
    else if (P.N > 0)
        hipLaunchKernelGGL(fwd_plain_fixed_kernel, dim3((unsigned)((P.N + 255) / 256)), dim3(256), 0, s, P, I,
                           (const Ctrl*)ctrl, acc);
    hipLaunchKernelGGL(guarded_convert_kernel, dim3(1024), dim3(256), 0, s, (const Ctrl*)ctrl,
                       (const unsigned long long*)acc, out, numel);
    return hipGetLastError() == hipSuccess ? PDS_OK : PDS_ERR_HIP;
}

}  // namespace

size_t binned_workspace_bytes(int64_t N, int32_t C, int32_t H, int32_t W, int32_t cap, int32_t mode,
                              const pds_tuning* t, bool ell_trivial, bool keep) {
    if (mode == PDS_MODE_PLAIN) return 0;
    const bool det = mode == PDS_MODE_DETERMINISTIC;
    Geom g;
    if (!ell_trivial || !make_geom(N, C, H, W, cap, mode, t, g)) {
        // photon-order paths: only the deterministic forward needs the int64 image and a control block
        if (keep || !det) return 0;
        return align256(sizeof(Ctrl)) + sizeof(uint64_t) * (size_t)C * H * W;
    }
    return make_layout(g, N, C, H, W, keep).end;
}

int binned_forward(const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning* t, float* out, void* ws,
                   size_t ws_bytes, bool keep, hipStream_t s) {
    const bool det = mode == PDS_MODE_DETERMINISTIC;
    Geom g;
    if (!binnable(I, P.N, mode, t, g)) {
        if (keep) {
            g_err = "this problem cannot be binned (kept bins unavailable)";
            return PDS_ERR_ARGUMENT;
        }
        return forward_unbinned(P, I, det, out, ws, ws_bytes, s);
    }
    const Layout L = make_layout(g, P.N, I.C, I.H, I.W, keep);
    if (ws == nullptr || ws_bytes < L.end) {
        g_err = "workspace too small";
        return PDS_ERR_WORKSPACE;
    }
    const Buffers B = buffers(ws, L);
    launch_build(P, I, g, B, keep, out, s);
    launch_forward_tiles(P, I, g, B, out, keep, s);
    return hipGetLastError() == hipSuccess ? PDS_OK : PDS_ERR_HIP;
}

int binned_backward(const float* grad, const PhotonArgs& P, const ImageArgs& I, int32_t mode, const pds_tuning* t,
                    const GradArgs& G, void* ws, size_t ws_bytes, bool reuse, hipStream_t s) {
    Geom g;
    if (!binnable(I, P.N, mode, t, g)) {
        g_err = "this problem cannot be binned";
        return PDS_ERR_ARGUMENT;
    }
    const Layout L = make_layout(g, P.N, I.C, I.H, I.W, /*keep=*/true);
    if (ws == nullptr || ws_bytes < L.end) {
        g_err = "workspace too small";
        return PDS_ERR_WORKSPACE;
    }
    const Buffers B = buffers(ws, L);
    // no image output: the scatter clears nothing
    if (!reuse) launch_build(P, I, g, B, /*keep=*/true, nullptr, s);
    // reuse: the forward of these bins also left every segment's walk order
    launch_backward_tiles(grad, P, I, g, B, G, reuse, s);
    return hipGetLastError() == hipSuccess ? PDS_OK : PDS_ERR_HIP;
}

int binned_stats(const void* ws, size_t ws_bytes, int64_t N, const ImageArgs& I, int32_t mode, const pds_tuning* t,
                 pds_bins_stats* out, hipStream_t s) {
    Geom g;
    if (!binnable(I, N, mode, t, g)) {
        g_err = "this problem has no bins";
        return PDS_ERR_ARGUMENT;
    }
    const Layout L = make_layout(g, N, I.C, I.H, I.W, /*keep=*/true);
    if (ws == nullptr || ws_bytes < L.end) {
        g_err = "bins buffer too small";
        return PDS_ERR_WORKSPACE;
    }
    Ctrl c;
    if (hipMemcpyAsync(&c, (const unsigned char*)ws + L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return PDS_ERR_HIP;
    out->records = (int64_t)c.total;
    out->capacity = g.cap_records;
    out->segments = (int32_t)c.nseg;
    out->overflow = (int32_t)c.overflow;
    out->tile = g.T;
    out->seg_photons = g.seg;
    out->bins = g.nbins;
    out->block_photons = (int32_t)g.per_block;
    out->reserved[0] = out->reserved[1] = 0;
    return PDS_OK;
}

bool binned_possible(const ImageArgs& I, int64_t N, int32_t mode, const pds_tuning* t) {
    Geom g;
    return binnable(I, N, mode, t, g);
}

}  // namespace pds
