// Shared pieces of the gather translation units: the gather object, the
// packed AVPL record layout, compensated sums, pixel loads and the guarded
// antiradiance decisions (cudaGather.cu:109-144, Shaders/Gather.frag:149-188).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "arc_internal.hpp"
#include "gather_math.hpp"

namespace arc {
// One view's moment cut, as a launch over many views reads it: the tree's
// arrays, the G-buffer, the pixel thresholds (estimate, image sum and the
// factors of Threshold), the cut's parameters, the outputs, and where the
// view's tiles start in the launch's list of tiles (and their cost slots).
struct CutView {
    const arc_bvh_node* nodes;
    const float4* crec;
    const float4* mom;
    const float4* nbox;  // per internal node: the box of its AVPLs' unit normals (lo, hi)
    const float4* gpos;
    const float4* gnrm;
    const float4* galb;
    const arc_material* mats;
    const float* th_est;
    const float* th_sum;
    float th_eps, th_floor, th_inv_npix;
    int W, H;
    float delta;
    int budget, criterion;
    float open2;
    int exact_count;
    float* tile_cost;
    float4* out_rad;
    unsigned long long* counters;
    int tile_base, ntiles;
};
}  // namespace arc

struct arc_gather {
    arc_ctx* ctx = nullptr;
    int width = 0, height = 0;
    arc::DevBuf mats;
    int num_mats = 0;
    bool general = false;  // some material has a specular lobe
    const float4* gpos = nullptr;
    const float4* gnrm = nullptr;
    const float4* galb = nullptr;
    float4* out_res = nullptr;
    float4* out_rad = nullptr;
    float4* out_anti = nullptr;
    arc::DevBuf int_rad, int_anti, int_res;
    arc::DevBuf rec, part, counters;
    arc::DevBuf stage;  // AVPL SoA staged from host memory
    // clustered antiradiance: compacted AVPL indices (+ block offsets and
    // count), cluster boxes, list lengths and lists
    arc::DevBuf anti_idx, cbox, ccount, clist;
    arc::DevBuf crec;  // error-bound cut: packed node records
    arc::DevBuf cest;  // error-bound cut: per-pixel estimates and their sum
    // moment estimator: per-internal-node records, and the FP64 prefix sums
    // over the sorted leaves (with per-block offsets) they are made from
    arc::DevBuf mom, mom_pre, mom_tot;
    arc::DevBuf mom_nbox;  // per internal node: the box of its AVPLs' unit normals
    bool mom_ready = false;  // mom holds the records of the current cut tree
    // moment cut: each 8x8 tile's cost in the last frame (VALU instructions,
    // estimated from its counts) and the launch order made from it, heaviest
    // tile first; tile_cost_n is the tile count the costs describe (0: none)
    arc::DevBuf tile_cost, tile_order;
    int tile_cost_n = 0;
    // moment cut: the kept nodes handed from the traversal to the
    // matrix-core pass, est_chunks chunks of 128 per tile (0, the default:
    // estimated in the traversal, measured faster), each tile's chain, its
    // queued count and the pass's order
    int est_chunks = 0;
    // launch shapes the results do not depend on: the moment cut's traversal
    // stack (64, 96 or 128 nodes) and the clustered antiradiance's list
    // entries filed per pass (768, 1536 or 3072)
    int cut_stack = 96, anti_chunk = 1536;
    arc::DevBuf est_entries, est_meta, est_head, est_queued, est_order;
    arc_cutnet* net = nullptr;  // learned cut: the predictor (borrowed)
    arc_bvh* bvh = nullptr;
    // per-frame phase events: frame slot, then 9 marks (start, upload done,
    // bvh done, radiance done, antiradiance done, end, start of the main
    // radiance kernel, start and end of the main antiradiance kernel)
    static constexpr int kRing = 64;
    static constexpr int kMarks = 9;
    hipEvent_t ev[kRing][kMarks] = {};
    // the clustered antiradiance of a cut frame runs on a stream of its own,
    // forked from and joined back to the context's stream
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int ring_next = 0, ring_count = 0;
    // one launch of the moment cut over many views (arc_batch): with
    // defer_cut, a frame stops where its cut kernel would be launched,
    // leaving the cut's descriptor and the combine's arguments here for
    // arc::gather_finish
    bool defer_cut = false, deferred = false;
    arc::CutView def_view{};
    int def_rad_splits = 0, def_anti_splits = 0;
    float4* def_rad = nullptr;
};


namespace arc {
namespace gd {

using namespace arc::gm;

constexpr int kThreads = 256;

// Workgroups of a launch are dealt to the 8 XCDs round robin (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  xcd_block maps the launch index
// to a logical block so that each XCD runs one contiguous range of logical
// blocks: neighbouring screen blocks, which read the same tree nodes and AVPL
// lists, then share an L2.  A bijection on [0, nblocks).
__device__ inline int xcd_block(int b, int nblocks) {
    constexpr int kXcds = 8;
    const int x = b % kXcds, k = b / kXcds;
    int off = 0;
    for (int y = 0; y < x; ++y) off += (nblocks - y + kXcds - 1) / kXcds;
    return off + k;
}
constexpr int kRecF4 = 5;  // float4 per packed AVPL record
constexpr int kRun = 64;   // pairs summed plainly before a compensated add
// Relative guard band of the FP32 antiradiance decisions (FP32 rounding of
// these quantities is below 1e-6 relative); pairs inside it are re-decided
// in FP64.
constexpr float kGuard = 1e-5f;

enum Counter {
    C_VISITED = 0, C_EVALUATED = 1, C_ANTI_VISITED = 2, C_ANTI_EVALUATED = 3, C_ERRORS = 4, C_BUDGET = 5,
    C_FALLBACK = 6, C_TESTED = 7, C_EXACT = 8, C_NUM = 9
};

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

__device__ inline f3 f3of(float4 v) { return f3{v.x, v.y, v.z}; }


// ----------------------------------------------------------- pixels ----
struct PixelF {
    float px, py, pz;  // position
    float nx, ny, nz;  // unit normal
    float br, bg, bb;  // diffuse / pi (Lambertian pixel BRDF)
    int mat;
};

__device__ inline PixelF load_pixel(const float4* gpos, const float4* gnrm, const float4* galb,
                                    const arc_material* mats, int pix) {
    PixelF P;
    const float4 p = gpos[pix], n = gnrm[pix];
    P.px = p.x; P.py = p.y; P.pz = p.z;
    const float s = rsq(n.x * n.x + n.y * n.y + n.z * n.z);
    P.nx = n.x * s; P.ny = n.y * s; P.nz = n.z * s;
    P.mat = (int)n.w;
    float dr, dg, db;
    if (galb) {
        const float4 a = galb[pix];
        dr = a.x; dg = a.y; db = a.z;
    } else {
        const arc_material& m = mats[mat_slot(P.mat)];
        dr = m.diffuse.x; dg = m.diffuse.y; db = m.diffuse.z;
    }
    P.br = kOneOverPi * dr; P.bg = kOneOverPi * dg; P.bb = kOneOverPi * db;
    return P;
}


// ------------------------------------------------------- decisions ----
// The antiradiance tests are discontinuous (inside / outside a disk or cone,
// sign tests of dot products).  A decision is made in FP32 when the FP32
// value is farther from its threshold than a guard that bounds the FP32
// rounding error (with a wide margin); pairs inside the guard band are
// re-decided in FP64 from the raw inputs with the reference's formulas, which
// is the decision the FP64 oracle makes.  Pairs in the band are rare except
// for coplanar pixel / AVPL pairs, whose signs the reference leaves to
// rounding noise.

// getAntiradiance's tests (cudaGather.cu:116-140) in FP64.
__device__ __attribute__((noinline)) inline bool disk_accept_f64(float4 p4, float4 n4, f3 a, f3 an, f3 w, float radius) {
    using V3 = V<double>;
    const V3 pos = mk<double>(p4.x, p4.y, p4.z), nrm = mk<double>(n4.x, n4.y, n4.z);
    const V3 apos = ld3<double>(a), anrm = ld3<double>(an), aw = ld3<double>(w);
    const double denom = dot(aw, nrm);
    if (denom == 0.0) return false;
    const double t = dot(pos - apos, nrm) / denom;
    if (t <= 0.0) return false;
    const V3 hit = apos + t * normalize(aw);
    if (length(hit - pos) >= (double)radius) return false;
    if (dot(normalize(hit - apos), aw) <= 0.0) return false;
    if (dot(nrm, aw) >= 0.0) return false;
    if (dot(normalize(pos - apos), anrm) >= 0.0) return false;
    return true;
}

// GetAntiradiance's tests (Shaders/Gather.frag:155-166) in FP64.
__device__ __attribute__((noinline)) inline bool cone_accept_f64(float4 p4, float4 n4, f3 a, f3 w, double cos_half) {
    using V3 = V<double>;
    const V3 pos = mk<double>(p4.x, p4.y, p4.z), nrm = mk<double>(n4.x, n4.y, n4.z);
    const V3 apos = ld3<double>(a), aw = ld3<double>(w);
    const V3 to_pos = normalize(pos - apos);
    const double c = dot(to_pos, normalize(aw));
    if (c <= (double)kConeEps) return false;
    if (dot(nrm, -to_pos) <= (double)kConeEps) return false;
    return c > cos_half;
}


struct AntiParams {
    float radius, r2;        // disk radius and its square
    float cos_half, cone_lo; // cone: cos(half-angle), lower edge of the guard band
    double cos_half_d;
};

__host__ inline AntiParams anti_params(float radius, double cos_half) {
    AntiParams a;
    a.radius = radius;
    a.r2 = radius * radius;
    a.cos_half = (float)cos_half;
    a.cone_lo = std::max(kConeEps, (float)cos_half) - kGuard;
    a.cos_half_d = cos_half;
    return a;
}

// Antiradiance weight of one (pixel, AVPL) pair with d = a - p, dd = |d|^2,
// inv = 1/|d|, cn = n_p . d and ca = -n_a . d: 0 or 1 for the disk model
// (the record already holds A / (pi r^2)), cos_p / |d|^2 for the cone model.
// FP32 decisions outside the guard band; FP64 re-decision inside it.
template <int MODEL>
__device__ inline float anti_weight(const PixelF& P, const float4& r0, const float4& r1, const float4& r2v,
                                    float guard_w, float dx, float dy, float dz, float dd, float inv, float cn,
                                    float ca, const AntiParams& ap, const float4* gpos, const float4* gnrm, int q,
                                    const AvplsDev& raw, int i) {
    if (MODEL == ARC_ANTIRAD_DISK) {
        // line a + t w hits the tangent plane at t = -cn / (w . n) > 0, within the
        // disk, on the far side of the AVPL's surface, pixel facing against w.
        // With h = d + t w_unit, |h| < r  <=>  |h denom|^2 < r^2 denom^2, and
        // h denom = d denom - cn w_unit: no division needed.
        const float denom = r2v.x * P.nx + r2v.y * P.ny + r2v.z * P.nz;
        const float dl = kGuard * (dd * inv);
        // Outside the guard bands the signs are certain, and a pair that
        // fails one of them contributes 0 in FP32 and in FP64 alike: skip the
        // rest (most pairs of a tile fail the same sign test together).
        if (!(denom < guard_w && cn > -dl && ca < dl)) return 0.f;
        const float hx = dx * denom - cn * r0.w;
        const float hy = dy * denom - cn * r1.w;
        const float hz = dz * denom - cn * r2v.w;
        const float h2 = hx * hx + hy * hy + hz * hz;
        const float dn2 = denom * denom;
        const float lim = ap.r2 * dn2;
        const float g = kGuard * (dd * dn2 + cn * cn);
        bool in = denom < 0.f && cn > 0.f && h2 < lim && ca < 0.f;
        const bool amb = h2 < lim + g &&
                         (!(fabsf(h2 - lim) > g) || !(fabsf(cn) > dl) || !(fabsf(ca) > dl) ||
                          !(fabsf(denom) > guard_w));
        if (amb) in = disk_accept_f64(gpos[q], gnrm[q], raw.pos[i], raw.nrm[i], raw.dir[i], ap.radius);
        return in ? 1.f : 0.f;
    } else {
        // cone of half-angle acos(cos_half) about w, cosine at the pixel / dist^2
        const float c = -(dx * r0.w + dy * r1.w + dz * r2v.w) * inv;
        const float cp = cn * inv;
        bool in = c > kConeEps && cp > kConeEps && c > ap.cos_half;
        const bool amb = c > ap.cone_lo && cp > kConeEps - kGuard &&
                         (!(fabsf(c - ap.cos_half) > kGuard) || !(fabsf(cp - kConeEps) > kGuard) ||
                          !(fabsf(c - kConeEps) > kGuard));
        if (amb) in = cone_accept_f64(gpos[q], gnrm[q], raw.pos[i], raw.dir[i], ap.cos_half_d);
        return in ? cp * (inv * inv) : 0.f;
    }
}

}  // namespace gd

// Error-bounded light cut (gather_cut.hip).
struct CutEBParams {
    float eps;          // errorBound
    int d0;             // cutEstimateDepth
    float delta;        // distance floor of the bounds
    int budget;         // cutBudget
    int per_tile;       // cutPerTile
    int criterion;      // cutCriterion
    float image_floor;  // cutImageFloor
    float confidence;   // cutConfidence (learned cut)
    int eval_depth;     // cutEvalDepth
    int exact_count;    // cutExactCount (moment estimator)
    float opening;      // cutOpening (ARC_CRITERION_EXPANSION)
};
// `mark` (may be null) is recorded right before the main traversal kernel.
int cut_error_bound(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                    hipEvent_t mark);
// The moment estimator's node records for this frame's cut tree (after
// bvh_build_cut): FP64 prefix sums over the sorted leaves, then one record
// per internal node.
int cut_moments_build(arc_gather* g, arc_bvh* b);
// The error-bound cut per 8x8 tile with the moment estimator.  With `defer`,
// everything up to the cut kernel is enqueued and the view's descriptor is
// written there instead of launching it (cut_moments_views launches it).
int cut_moments(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                hipEvent_t mark, CutView* defer = nullptr);
// One launch of the moment cut over the tiles of many views (device array
// `views`, nitems tiles in all), heaviest tile first by the costs of the
// last launch when `ordered`; costs and order hold nitems each.
int cut_moments_views(arc_ctx* ctx, hipStream_t stream, const CutView* views, int nviews, int nitems,
                      float* costs, int* order, bool ordered);
// The combine of a frame run with defer_cut, once its cut has run.
int gather_finish(arc_gather* g);
// The learned cut: per tile, refine / keep from g->net, the error-bound
// criterion where the prediction is not confident.
int cut_learned(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                hipEvent_t mark);
// Training rows with exact labels (arc_gather_cut_samples); count: one device int.
int cut_samples(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float* X, float* y, int capacity, uint64_t seed,
                float keep, int* count);

// Clustered antiradiance (gather_anti.hip).  `extent` (the scene's extent)
// scales the margins of the cone model's conservative tests; `mark` (may be
// null) is recorded right before the per-pixel kernel.
int anti_clustered(arc_gather* g, const AvplsDev& a, int model, float3 cam, const gd::AntiParams& ap,
                   double half_angle, float extent, float* part, unsigned long long* counters, hipEvent_t mark);

}  // namespace arc
