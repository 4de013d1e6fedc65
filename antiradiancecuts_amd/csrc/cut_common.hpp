// Pieces shared by the light-cut kernels (gather_cut.hip, cut_moments.hip):
// the node records the cuts read, the exact G of a leaf and the G~ of a
// node's centroid, the Lightcuts contribution bound, compensated per-pixel
// sums, the tile-to-pixel mapping and the per-pixel thresholds.
#pragma once

#include <cmath>

#include "gather_common.hpp"

namespace arc {
namespace cut {

using arc::f3;
using arc::kCrec;
using namespace arc::gm;
using namespace arc::gd;

constexpr int kCutThreads = 256;

struct CutEB {
    const arc_bvh_node* nodes;  // n - 1 internal nodes
    const float4* crec;         // arc::kCrec float4 per slot (arc::write_cut_record)
    int n;
};

// G~(p, node) (cudaGather.cu:45-57 with the node's moments); for a leaf the
// exact G of the AVPL.
__device__ inline float geom_est(const PixelF& P, const float4& r0, const float4& r1) {
    const float dx = r0.x - P.px, dy = r0.y - P.py, dz = r0.z - P.pz;
    const float dd = dx * dx + dy * dy + dz * dz;
    // coincident points (closer than 1e-15): no contribution, not 0 * inf
    const float inv = dd > 1e-30f ? __builtin_amdgcn_rsqf(dd) : 0.f;
    const float c1 = fminf(fmaxf((P.nx * dx + P.ny * dy + P.nz * dz) * inv, 0.f), 1.f);
    // r1.w >= 0, so the sum is >= 0 and clamping it to [0, 1] is min(sum, 1):
    // written as a clamp, the add carries it as an output modifier
    const float c2 =
        fminf(fmaxf(r1.w + fmaxf(-(r1.x * dx + r1.y * dy + r1.z * dz) * inv, 0.f), 0.f), 1.f);
    return c1 * c2 * (inv * inv);
}

__device__ inline float node_bound(const PixelF& P, const arc_bvh_node& b, float ilum, float delta2) {
    const float ex = fmaxf(fmaxf(b.bbMin.x - P.px, P.px - b.bbMax.x), 0.f);
    const float ey = fmaxf(fmaxf(b.bbMin.y - P.py, P.py - b.bbMax.y), 0.f);
    const float ez = fmaxf(fmaxf(b.bbMin.z - P.pz, P.pz - b.bbMax.z), 0.f);
    const float dmin2 = fmaxf(ex * ex + ey * ey + ez * ez, delta2);
    const float hmax = P.nx * (P.nx > 0.f ? b.bbMax.x - P.px : b.bbMin.x - P.px) +
                       P.ny * (P.ny > 0.f ? b.bbMax.y - P.py : b.bbMin.y - P.py) +
                       P.nz * (P.nz > 0.f ? b.bbMax.z - P.pz : b.bbMin.z - P.pz);
    // 0 when the node lies behind the pixel's tangent plane
    return hmax > 0.f ? ilum * fminf(1.f, hmax * __builtin_amdgcn_rsqf(dmin2)) * __builtin_amdgcn_rcpf(dmin2) : 0.f;
}

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

struct Acc {
    Kahan k[3];
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    int run = 0;
    // one term of the plain run, without a flush test: a kept node's
    // estimates then stay one straight run of code (a flush test between them
    // costs register moves at the join); the caller flushes once per kept
    // node, or per batch of kept nodes (maybe_flush), so a plain run can
    // exceed kRun by that node's or batch's extra terms
    __device__ inline void add_run(const float4& I, float g) {
        s0 += I.x * g;
        s1 += I.y * g;
        s2 += I.z * g;
        ++run;
    }
    __device__ inline void maybe_flush() {
        if (run >= kRun) flush();
    }
    __device__ inline void flush() {
        k[0].add(s0); k[1].add(s1); k[2].add(s2);
        s0 = s1 = s2 = 0.f;
        run = 0;
    }
};

// Pixel of lane `lane` of wave `wave` in a 16x16 block: four 8x8 tiles.
// The launch is a 1D grid of gx x gy such blocks, taken in XCD-contiguous order.
__device__ inline int2 tile_pixel(int wave, int lane, int gx, int gy) {
    const int b = xcd_block(blockIdx.x, gx * gy);
    return make_int2((b % gx) * 16 + (wave & 1) * 8 + (lane & 7), (b / gx) * 16 + (wave >> 1) * 8 + (lane >> 3));
}

struct Threshold {
    const float* est;  // per-pixel estimate of the depth-d0 cut
    const float* sum;  // its sum over the image
    float eps, floor_frac, inv_npix;
    // eps * max(estimate, floor_frac * image-mean estimate)
    __device__ inline float at(int q) const {
        return fmaxf(eps * fmaxf(est[q], floor_frac * sum[0] * inv_npix), 1e-30f);
    }
};

// Per-pixel estimates of the depth-d0 cut and the thresholds (the node
// records come with the tree, bvh_build_cut); defined in gather_cut.hip.
int prepare_cut(arc_gather* g, arc_bvh* b, const CutEBParams& cp, CutEB& t, Threshold& th);

}  // namespace cut
}  // namespace arc
