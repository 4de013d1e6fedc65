// Error-bounded light-cut radiance on gfx950 (ARC_CUT_ERROR_BOUND).
//
// The reference refines a cluster while its solid angle exceeds a threshold
// (kernel_bvh, cudaGather.cu:405-414).  This mode uses the Lightcuts
// criterion instead (Walter et al., SIGGRAPH 2005): a cluster is kept when an
// upper bound of its contribution is at most errorBound times the pixel's
// radiance estimate, and refined otherwise.
//
//   cluster estimate  I * G~   with I the sum of L * diffuse / pi over the
//                     node's AVPLs (Lambertian AVPLs; the pixel BRDF is
//                     applied once per pixel) and G~ evaluated at the
//                     w-weighted centroid c of the node (w = mean of the
//                     three channels of L * diffuse / pi):
//                       G~ = cos_p(c) * cos~ / |c - p|^2,
//                       cos~ = (1 - |m|) / 4 + max(0, -m . (p - c)^),
//                     m the w-weighted mean unit normal.  cos~ is exact for
//                     one AVPL or a node whose normals agree (|m| = 1) and
//                     matches the average clamped cosine of isotropic normals
//                     (|m| = 0, mean 1/4); the reference's random
//                     representative (bvh.cu:726-744) is unbiased but has a
//                     spread of the order of the node's contribution when
//                     normals vary, which the deterministic moment estimate
//                     removes.
//   bound             max(I) * min(1, h_max / d_min) / d_min^2 over the node's
//                     box: d_min is the distance from p to the box, floored at
//                     delta (1e-3 of the scene extent), h_max the largest
//                     height of the box over the pixel's tangent plane (<= 0:
//                     the whole node is behind the pixel and contributes
//                     exactly 0); the AVPL-side cosine is bounded by 1.
//   threshold         errorBound * max(e, cutImageFloor * mean of e over the
//                     image), e the pixel's estimate from the cut made of the
//                     nodes at depth d0 (k_cut_estimate).  The floor keeps
//                     pixels that see almost nothing from refining to the
//                     leaves: the image error is an L2 norm, so a dark pixel
//                     needs no finer absolute error than an average one, and
//                     the longest cut sets the kernel's time.
//
// Two traversals share the node records:
//   k_cut_eb_px   one pixel per lane, its own depth-first cut (the reference's
//                 per-pixel traversal), stack in LDS;
//   k_cut_eb_tile one 8x8 tile per wave: one depth-first cut for the tile,
//                 refined when any lane's bound exceeds its threshold (ballot),
//                 wave-uniform stack and scalar node loads.  It pays off where
//                 neighbouring pixels see the AVPLs alike (smooth surfaces far
//                 from the AVPLs); with AVPLs around the pixels the tile cut
//                 is the union of 64 different near fields.
#include <algorithm>
#include <cmath>

#include "cut_common.hpp"
#include "cutnet.hpp"

using arc::f3;
using arc::kCrec;
using namespace arc::gm;
using namespace arc::gd;
namespace net = arc::net;
using namespace arc::cut;

namespace {

constexpr int kTileStack = 96;  // per-wave stack of the tile traversal
constexpr int kPxStack = 48;    // per-lane stack of the pixel traversal (LDS)
constexpr int kEstStack = 48;   // per-wave stack of the depth-d0 estimate cut (d0 <= 40)

// Estimated error of keeping the node instead of its AVPLs (criterion 1): the
// contribution bound seen from the centroid, max(I) cos_max / (d_c - s)^2
// with s the half-diagonal of the box, times the estimator's relative error:
// (s / d_c)^2 for the position (centroid, second order) plus
// 1.2 (1 - |m|) / sqrt(count) for the spread of the clamped cosines of
// differing normals (their mean over `count` normals has that relative
// spread).  A pixel within 2 s of the centroid always refines.
__device__ inline float node_error(const PixelF& P, const arc_bvh_node& b, const float4& r0, const float4& r2,
                                   const float4& r3, float delta) {
    // the box corners as values first: selecting between them per lane then
    // costs a v_cndmask instead of a per-lane (vector) load
    const float x0 = b.bbMin.x, y0 = b.bbMin.y, z0 = b.bbMin.z;
    const float x1 = b.bbMax.x, y1 = b.bbMax.y, z1 = b.bbMax.z;
    const float dx = r0.x - P.px, dy = r0.y - P.py, dz = r0.z - P.pz;
    const float dc2 = dx * dx + dy * dy + dz * dz;
    const float hmax = P.nx * ((P.nx > 0.f ? x1 : x0) - P.px) + P.ny * ((P.ny > 0.f ? y1 : y0) - P.py) +
                       P.nz * ((P.nz > 0.f ? z1 : z0) - P.pz);
    // branch-free: a heuristic, hardware square root and reciprocal (about
    // 1 ulp) suffice
    const float near = fmaxf(__builtin_amdgcn_sqrtf(dc2) - r3.y, delta);
    const float inv_near = __builtin_amdgcn_rcpf(near);
    // clamped at 0 as well: an output modifier of the multiply (the result
    // for hmax <= 0 is 0 below either way)
    const float cb = fminf(fmaxf(hmax * inv_near, 0.f), 1.f);
    const float rel = r3.x * __builtin_amdgcn_rcpf(dc2) + r2.w;
    const float e = r0.w * cb * (inv_near * inv_near) * rel;
    // behind the pixel's tangent plane: exactly 0; within 2 s of the centroid: refine
    return dc2 <= 4.f * r3.x ? INFINITY : (hmax > 0.f ? e : 0.f);
}

// A kept node's contribution to the pixel: its own estimate, or the sum of
// its descendants' estimates D levels down (leaves where the subtree ends
// sooner).  A finer sum for the same node test: the descendants' centroids
// sit nearer their AVPLs.  Returns the estimates added.
template <int D>
__device__ inline int add_desc(const CutEB& t, const PixelF& P, int v, Acc& acc) {
    if (D == 0 || v >= 0) {
        const float4* r = t.crec + kCrec * (size_t)(v >= 0 ? v : t.n + (-(v + 1)));
        acc.add_run(r[2], geom_est(P, r[0], r[1]));
        return 1;
    }
    const arc_bvh_node& b = t.nodes[-(v + 1)];
    const int l = b.left, r = b.right;
    return add_desc<(D > 0 ? D - 1 : 0)>(t, P, l, acc) + add_desc<(D > 0 ? D - 1 : 0)>(t, P, r, acc);
}

constexpr int kMaxEvalDepth = 4;

__device__ inline int add_kept(const CutEB& t, const PixelF& P, int node, int depth, Acc& acc) {
    switch (depth) {
        case 1: return add_desc<1>(t, P, node, acc);
        case 2: return add_desc<2>(t, P, node, acc);
        case 3: return add_desc<3>(t, P, node, acc);
        case 4: return add_desc<4>(t, P, node, acc);
        default: return add_desc<0>(t, P, node, acc);
    }
}

__global__ void k_cut_zero(float* __restrict__ sum) {
    if (threadIdx.x == 0) sum[0] = 0.f;
}

// Phase 1: the pixel's estimate I * G~ summed over the cut made of every node
// at depth d0 (and the leaves above it).  The cut is the same for every
// pixel, so the wave walks it uniformly; each lane sums its own pixel.
__global__ __launch_bounds__(kCutThreads) void k_cut_estimate(CutEB t, const float4* __restrict__ gpos,
                                                              const float4* __restrict__ gnrm,
                                                              const float4* __restrict__ galb,
                                                              const arc_material* __restrict__ mats, int W, int H,
                                                              int d0, float* __restrict__ est,
                                                              float* __restrict__ sum) {
    __shared__ int stk[kCutThreads / 64][kEstStack];
    __shared__ float red[kCutThreads / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int2 xy = tile_pixel(wave, lane, (W + 15) / 16, (H + 15) / 16);
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    int* S = stk[wave];
    const int n = t.n;
    float e = 0.f;
    int sp = 1;
    S[0] = -64;  // node * 64 + depth
    while (sp > 0) {
        const int v = uniform(S[--sp]);
        const int node = v >> 6, dep = v & 63;
        if (node >= 0 || dep >= d0) {
            const float4* r = t.crec + kCrec * (size_t)(node >= 0 ? node : n + (-(node + 1)));
            e += r[0].w * geom_est(P, r[0], r[1]);
            continue;
        }
        const arc_bvh_node& b = t.nodes[-(node + 1)];
        S[sp] = b.left * 64 + dep + 1;
        S[sp + 1] = b.right * 64 + dep + 1;
        sp += 2;
    }
    if (live) est[q] = e;
    float w = live ? e : 0.f;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off);
    if (lane == 0) red[wave] = w;
    __syncthreads();
    if (tid == 0) atomicAdd(sum, red[0] + red[1] + red[2] + red[3]);
}

// The same estimate with the depth-d0 cut listed per block.  The walk above
// is a chain of dependent loads per wave (stack, node, children); here each
// block first expands the frontier level by level (a node stays when it is
// a leaf or at depth d0; wave ballots give each entry its place), stages the
// cut's (r0, r1) records in LDS right to left, the order in which the walk
// sums them, and then every lane streams them: both kernels give the same
// sums.  The frontier holds at most 2^d0 <= kCutThreads entries.
constexpr int kEstListDepth = 8;
static_assert((1 << kEstListDepth) <= kCutThreads, "one frontier entry per thread");

__global__ __launch_bounds__(kCutThreads) void k_cut_estimate_list(CutEB t, int d0, const float4* __restrict__ gpos,
                                                                   const float4* __restrict__ gnrm,
                                                                   const float4* __restrict__ galb,
                                                                   const arc_material* __restrict__ mats, int W,
                                                                   int H, float* __restrict__ est,
                                                                   float* __restrict__ sum) {
    extern __shared__ float4 cut[];  // (r0, r1) per cut node: 2 x min(2^d0, n) entries
    __shared__ int fr[2][kCutThreads];
    __shared__ int wsum[kCutThreads / 64];
    __shared__ float red[kCutThreads / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint64_t below = (1ull << lane) - 1;
    int cur = 0, m = 1;
    if (tid == 0) fr[0][0] = -1;  // the root
    __syncthreads();
    for (int dep = 0; dep < d0; ++dep) {  // m <= 2^dep <= kCutThreads / 2
        const bool act = tid < m;
        const int v = act ? fr[cur][tid] : 0;
        const bool inner = act && v < 0;
        const uint64_t ba = __ballot(act), bi = __ballot(inner);
        if (lane == 0) wsum[wave] = __popcll(ba) + __popcll(bi);
        __syncthreads();
        int pos = __popcll(ba & below) + __popcll(bi & below), next = 0;
        for (int w = 0; w < kCutThreads / 64; ++w) {
            pos += w < wave ? wsum[w] : 0;
            next += wsum[w];
        }
        if (inner) {
            const arc_bvh_node& b = t.nodes[-(v + 1)];
            fr[1 - cur][pos] = b.left;
            fr[1 - cur][pos + 1] = b.right;
        } else if (act) {
            fr[1 - cur][pos] = v;
        }
        __syncthreads();  // the new frontier is complete and wsum is read
        cur = 1 - cur;
        m = next;
    }
    for (int k = tid; k < m; k += kCutThreads) {
        const int v = fr[cur][m - 1 - k];
        const float4* r = t.crec + kCrec * (size_t)(v >= 0 ? v : t.n + (-(v + 1)));
        cut[2 * k] = r[0];
        cut[2 * k + 1] = r[1];
    }
    __syncthreads();
    const int2 xy = tile_pixel(wave, lane, (W + 15) / 16, (H + 15) / 16);
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    float e = 0.f;
    for (int k = 0; k < m; ++k) {
        const float4 r0 = cut[2 * k], r1 = cut[2 * k + 1];
        e += r0.w * geom_est(P, r0, r1);
    }
    if (live) est[q] = e;
    float w = live ? e : 0.f;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off);
    if (lane == 0) red[wave] = w;
    __syncthreads();
    if (tid == 0) atomicAdd(sum, red[0] + red[1] + red[2] + red[3]);
}

// ------------------------------------------------------------ per pixel --
__global__ __launch_bounds__(kCutThreads) void k_cut_eb_px(CutEB t, const float4* __restrict__ gpos,
                                                           const float4* __restrict__ gnrm,
                                                           const float4* __restrict__ galb,
                                                           const arc_material* __restrict__ mats, int W, int H,
                                                           Threshold th, float delta, int budget, int criterion,
                                                           int eval_depth, float4* __restrict__ out_rad,
                                                           unsigned long long* __restrict__ counters) {
    __shared__ int stk[kPxStack * kCutThreads];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int2 xy = tile_pixel(wave, lane, (W + 15) / 16, (H + 15) / 16);
    const bool live = xy.x < W && xy.y < H;
    unsigned long long nvis = 0, nev = 0, nerr = 0, nbud = 0, ntest = 0;
    if (live) {
        const int q = xy.y * W + xy.x;
        const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
        const int n = t.n;
        int* S = stk + tid;  // entry j at S[j * kCutThreads]
        const float thr = th.at(q);
        // depth first: the right child waits on the stack while the left one
        // is processed at once
        Acc acc;
        int sp = 0;
        int node = -1;
        for (;;) {
            int slot = node;
            if (node < 0) {
                const int k = -(node + 1);
                const arc_bvh_node& b = t.nodes[k];
                slot = n + k;
                const float4* r = t.crec + kCrec * (size_t)slot;
                const float crit = criterion ? node_error(P, b, r[0], r[2], r[3], delta)
                                             : node_bound(P, b, r[0].w, delta * delta);
                ++ntest;
                if (crit > thr) {
                    if ((int)nev >= budget) {
                        ++nbud;  // cut budget spent: keep the cluster
                    } else if (sp < kPxStack) {
                        S[(sp++) * kCutThreads] = b.right;
                        node = b.left;
                        ++nvis;
                        continue;
                    } else {
                        ++nerr;  // stack full: keep the cluster
                    }
                }
            }
            (void)slot;
            nev += add_kept(t, P, node, eval_depth, acc);
            acc.maybe_flush();
            if (sp == 0) break;
            node = S[(--sp) * kCutThreads];
        }
        acc.flush();
        // with counters, w carries the pixel's cut size (kernel_bvh's debug output role)
        out_rad[q] = make_float4(acc.k[0].s * P.br, acc.k[1].s * P.bg, acc.k[2].s * P.bb,
                                 counters ? (float)nev : 1.f);
    }
    if (counters) {
        for (int off = 32; off > 0; off >>= 1) {
            nvis += __shfl_down(nvis, off);
            nev += __shfl_down(nev, off);
            nerr += __shfl_down(nerr, off);
            nbud += __shfl_down(nbud, off);
            ntest += __shfl_down(ntest, off);
        }
        if (lane == 0) {
            atomicAdd(&counters[C_TESTED], ntest);
            atomicAdd(&counters[C_VISITED], nvis);
            atomicAdd(&counters[C_EVALUATED], nev);
            atomicAdd(&counters[C_ERRORS], nerr);
            atomicAdd(&counters[C_BUDGET], nbud);
        }
    }
}

// ------------------------------------------------------------- per tile --
// One depth-first cut per 8x8 tile (one wave): a node is refined when any of
// the tile's pixels asks for it (ballot).  The refined node's left child is
// taken next without a round trip through the stack; its right child waits
// there.  Node indices, stack and records are wave-uniform (scalar loads).
template <int D>
__global__ __launch_bounds__(kCutThreads) void k_cut_eb_tile(CutEB t, const float4* __restrict__ gpos,
                                                             const float4* __restrict__ gnrm,
                                                             const float4* __restrict__ galb,
                                                             const arc_material* __restrict__ mats, int W, int H,
                                                             Threshold th, float delta, int budget, int criterion,
                                                             float4* __restrict__ out_rad,
                                                             unsigned long long* __restrict__ counters) {
    __shared__ int stk[kCutThreads / 64][kTileStack];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int2 xy = tile_pixel(wave, lane, (W + 15) / 16, (H + 15) / 16);
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    int* S = stk[wave];
    const int n = t.n;
    const float thr = live ? th.at(q) : INFINITY;
    const float delta2 = delta * delta;
    Acc acc;
    unsigned nvis = 0, nev = 0, nerr = 0, nbud = 0, ntest = 0;  // wave-uniform: scalar registers
    int sp = 0;
    int node = -1;
    for (;;) {
        int slot = node;
        if (node < 0) {
            const int k = -(node + 1);
            const arc_bvh_node& b = t.nodes[k];
            slot = n + k;
            const float4* r = t.crec + kCrec * (size_t)slot;
            const float crit = criterion ? node_error(P, b, r[0], r[2], r[3], delta)
                                         : node_bound(P, b, r[0].w, delta2);
            ++ntest;
            if (__ballot(crit > thr)) {
                if ((int)nev >= budget) {
                    ++nbud;  // cut budget spent: keep the cluster
                } else if (sp < kTileStack) {
                    S[sp++] = b.right;
                    node = uniform(b.left);
                    ++nvis;
                    continue;
                } else {
                    ++nerr;  // stack full: keep the cluster
                }
            }
        }
        (void)slot;
        nev += add_desc<D>(t, P, node, acc);
        acc.maybe_flush();
        if (sp == 0) break;
        node = uniform(S[--sp]);
    }
    acc.flush();
    const int alive = __popcll(__ballot(live));  // all lanes take part in the ballot
    if (counters && lane == 0) {
        atomicAdd(&counters[C_TESTED], (unsigned long long)ntest * alive);
        atomicAdd(&counters[C_VISITED], (unsigned long long)nvis * alive);
        atomicAdd(&counters[C_EVALUATED], (unsigned long long)nev * alive);
        atomicAdd(&counters[C_ERRORS], (unsigned long long)nerr);
        atomicAdd(&counters[C_BUDGET], (unsigned long long)nbud);
    }
    if (live)
        out_rad[q] = make_float4(acc.k[0].s * P.br, acc.k[1].s * P.bg, acc.k[2].s * P.bb,
                                 counters ? (float)nev : 1.f);
}

// ------------------------------------------------------- learned cut --
// The learned cut decides per (8x8 tile, node) from features seen at the
// tile's centre; the same feature function serves the labelled traversal
// that makes training rows and the inference traversal.

constexpr int kLearnStack = 1024;  // per-wave stack: a batch pops 16 nodes and may push 32
constexpr int kLearnBatch = 16;   // nodes decided together: one MFMA sample tile

__device__ inline float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ inline float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}
__device__ inline float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

// Wave-uniform statistics of a tile's live pixels.
struct TileStats {
    PixelF c;         // centre: mean position, normalised mean normal
    float nspread;    // 1 - |mean normal|
    float radius;     // largest distance of a pixel from the centre
    float thr_min, thr_max;
    float live_frac;
};

__device__ inline TileStats tile_stats(const PixelF& P, bool live, float thr) {
    TileStats T;
    const float cnt = wave_sum(live ? 1.f : 0.f);
    T.live_frac = cnt * (1.f / 64.f);
    const float ic = cnt > 0.f ? 1.f / cnt : 0.f;
    T.c.px = wave_sum(live ? P.px : 0.f) * ic;
    T.c.py = wave_sum(live ? P.py : 0.f) * ic;
    T.c.pz = wave_sum(live ? P.pz : 0.f) * ic;
    const float mx = wave_sum(live ? P.nx : 0.f) * ic, my = wave_sum(live ? P.ny : 0.f) * ic,
                mz = wave_sum(live ? P.nz : 0.f) * ic;
    const float ml = sqrtf(mx * mx + my * my + mz * mz);
    T.nspread = 1.f - fminf(ml, 1.f);
    const float im = ml > 0.f ? 1.f / ml : 0.f;
    T.c.nx = mx * im;
    T.c.ny = my * im;
    T.c.nz = ml > 0.f ? mz * im : 1.f;
    T.c.br = T.c.bg = T.c.bb = 0.f;
    T.c.mat = 0;
    const float dx = P.px - T.c.px, dy = P.py - T.c.py, dz = P.pz - T.c.pz;
    T.radius = wave_max(live ? sqrtf(dx * dx + dy * dy + dz * dz) : 0.f);
    T.thr_min = wave_min(live ? thr : INFINITY);
    T.thr_max = wave_max(live ? thr : 0.f);
    // every lane holds the same values: keep them in scalar registers
    T.c.px = __builtin_amdgcn_readfirstlane(T.c.px);
    T.c.py = __builtin_amdgcn_readfirstlane(T.c.py);
    T.c.pz = __builtin_amdgcn_readfirstlane(T.c.pz);
    T.c.nx = __builtin_amdgcn_readfirstlane(T.c.nx);
    T.c.ny = __builtin_amdgcn_readfirstlane(T.c.ny);
    T.c.nz = __builtin_amdgcn_readfirstlane(T.c.nz);
    T.nspread = __builtin_amdgcn_readfirstlane(T.nspread);
    T.radius = __builtin_amdgcn_readfirstlane(T.radius);
    T.thr_min = __builtin_amdgcn_readfirstlane(T.thr_min);
    T.thr_max = __builtin_amdgcn_readfirstlane(T.thr_max);
    T.live_frac = __builtin_amdgcn_readfirstlane(T.live_frac);
    return T;
}

// One corner pixel of the tile per lane group (lanes 16 g .. 16 g + 15):
// pixels (0,0), (7,0), (0,7), (7,7) of the 8x8 tile; a dead corner has an
// infinite threshold.
struct Corner {
    PixelF p;
    float thr;
};

__device__ inline Corner tile_corner(const PixelF& P, float thr, int g) {
    const int cl = g == 0 ? 0 : g == 1 ? 7 : g == 2 ? 56 : 63;
    Corner c;
    c.p.px = __shfl(P.px, cl);
    c.p.py = __shfl(P.py, cl);
    c.p.pz = __shfl(P.pz, cl);
    c.p.nx = __shfl(P.nx, cl);
    c.p.ny = __shfl(P.ny, cl);
    c.p.nz = __shfl(P.nz, cl);
    c.p.br = c.p.bg = c.p.bb = 0.f;
    c.p.mat = 0;
    c.thr = __shfl(thr, cl);
    return c;
}

__device__ inline float lg2(float x) { return __builtin_amdgcn_logf(fmaxf(x, 1e-30f)); }
__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// The 24 features of (tile, internal node).  Scale-free: distances enter as
// ratios, intensities relative to the tile's most demanding threshold.  Two
// of them are the error-bound criterion's terms at the tile's four corners
// (one corner per lane group, the largest over the groups), so all 64 lanes
// must call this, each group with the same node.
__device__ inline void cut_features(const TileStats& T, const Corner& K, const arc_bvh_node& b, const float4& r0,
                                    const float4& r1, const float4& r2, const float4& r3, float eps, float delta,
                                    float f[net::kIn]) {
    const PixelF& C = T.c;
    const float dx = r0.x - C.px, dy = r0.y - C.py, dz = r0.z - C.pz;
    const float dc2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-30f);
    const float inv = __builtin_amdgcn_rsqf(dc2), dc = dc2 * inv;
    const float ux = dx * inv, uy = dy * inv, uz = dz * inv;
    const float c1 = C.nx * ux + C.ny * uy + C.nz * uz;
    const float c2 = -(r1.x * ux + r1.y * uy + r1.z * uz);
    const float hmax = C.nx * (C.nx > 0.f ? b.bbMax.x - C.px : b.bbMin.x - C.px) +
                       C.ny * (C.ny > 0.f ? b.bbMax.y - C.py : b.bbMin.y - C.py) +
                       C.nz * (C.nz > 0.f ? b.bbMax.z - C.pz : b.bbMin.z - C.pz);
    const float est = r0.w * geom_est(C, r0, r1);
    const float err = node_error(C, b, r0, r2, r3, delta);
    const float bnd = node_bound(C, b, r0.w, delta * delta);
    const float itm = 1.f / fmaxf(T.thr_min, 1e-30f);
    f[0] = fminf(r3.y * inv, 8.f);
    f[1] = fminf(T.radius * inv, 8.f);
    f[2] = c1;
    f[3] = c2;
    f[4] = 1.f - 4.f * r1.w;
    f[5] = T.nspread;
    f[6] = lg2(r3.z) * (1.f / 16.f);
    f[7] = clampf(lg2(est * itm) * (1.f / 16.f), -4.f, 4.f);
    f[8] = err == INFINITY ? 4.f : clampf(lg2(err * itm) * (1.f / 16.f), -4.f, 4.f);
    f[9] = clampf(lg2(bnd * itm) * (1.f / 16.f), -4.f, 4.f);
    f[10] = clampf(hmax * inv, -4.f, 4.f);
    f[11] = clampf(lg2(dc2 / (4.f * r3.x + 1e-30f)) * 0.125f, -4.f, 4.f);
    f[12] = r2.w;
    f[13] = lg2(T.thr_max * itm) * 0.125f;
    f[14] = lg2(eps) * 0.125f;
    f[15] = T.live_frac;
    f[16] = clampf(lg2(dc / delta) * (1.f / 16.f), -4.f, 4.f);
    f[17] = C.nx * r1.x + C.ny * r1.y + C.nz * r1.z;
    f[18] = (r2.x + r2.y + r2.z) / (3.f * r0.w + 1e-30f);
    f[20] = clampf(c1, 0.f, 1.f);
    f[21] = fminf(1.f, r1.w + fmaxf(c2, 0.f));
    f[22] = hmax > 0.f ? fminf(1.f, hmax / fmaxf(dc - r3.y, delta)) : 0.f;
    // the criterion at this group's corner, then the largest over the corners
    const float ek = node_error(K.p, b, r0, r2, r3, delta);
    const float bk = node_bound(K.p, b, r0.w, delta * delta);
    float fe = ek == INFINITY ? 4.f : clampf(lg2(ek / K.thr) * (1.f / 16.f), -4.f, 4.f);
    float fb = clampf(lg2(bk / K.thr) * (1.f / 16.f), -4.f, 4.f);
    fe = fmaxf(fe, __shfl_xor(fe, 16));
    fe = fmaxf(fe, __shfl_xor(fe, 32));
    fb = fmaxf(fb, __shfl_xor(fb, 16));
    fb = fmaxf(fb, __shfl_xor(fb, 32));
    f[19] = fe;
    f[23] = fb;
}

__device__ inline uint32_t mix32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

struct LabelOut {
    float* X;
    float* y;
    int* count;
    int capacity;
    uint64_t seed;
    uint32_t keep;  // keep a row when its hash is below this
};

// The labelled traversal: one depth-first cut per tile whose refine
// decisions are the labels.  A node's label is "refine" when some live pixel
// of the tile sees the node's estimate I * G~ off the exact sum of the
// node's leaves by more than the pixel's threshold (largest channel).
__global__ __launch_bounds__(kCutThreads) void k_cut_label_tile(CutEB t, const int* __restrict__ ranges,
                                                                const float4* __restrict__ gpos,
                                                                const float4* __restrict__ gnrm,
                                                                const float4* __restrict__ galb,
                                                                const arc_material* __restrict__ mats, int W, int H,
                                                                Threshold th, float eps, float delta, LabelOut o) {
    __shared__ int stk[kCutThreads / 64][kTileStack];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16;
    const int2 xy = tile_pixel(wave, lane, gx, gy);
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    const float thr = live ? th.at(q) : INFINITY;
    const TileStats T = tile_stats(P, live, thr);
    const Corner K = tile_corner(P, thr, lane >> 4);
    if (T.live_frac == 0.f || t.n < 2) return;  // wave-uniform
    const uint64_t tile = (uint64_t)xcd_block(blockIdx.x, gx * gy) * 4 + wave;
    int* S = stk[wave];
    const int n = t.n;
    int sp = 0;
    int k = 0;  // internal node
    for (;;) {
        const arc_bvh_node& b = t.nodes[k];
        const float4* r = t.crec + kCrec * (size_t)(n + k);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
        float e0 = 0.f, e1 = 0.f, e2 = 0.f;
        const int first = uniform(ranges[2 * k]), last = uniform(ranges[2 * k + 1]);
        for (int s = first; s <= last; ++s) {
            const float4* lr = t.crec + kCrec * (size_t)s;
            const float ge = geom_est(P, lr[0], lr[1]);
            const float4 li = lr[2];
            e0 += li.x * ge;
            e1 += li.y * ge;
            e2 += li.z * ge;
        }
        const float ge = geom_est(P, r0, r1);
        const float err = fmaxf(fabsf(e0 - r2.x * ge), fmaxf(fabsf(e1 - r2.y * ge), fabsf(e2 - r2.z * ge)));
        const bool refine = __ballot(live && err > thr) != 0;
        if (mix32(o.seed ^ (tile * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)k << 1)) < o.keep) {
            float f[net::kIn];
            cut_features(T, K, b, r0, r1, r2, r3, eps, delta, f);
            if (lane == 0) {
                const int row = atomicAdd(o.count, 1);
                if (row < o.capacity) {
                    float* x = o.X + (size_t)row * net::kIn;
#pragma unroll
                    for (int j = 0; j < net::kIn; ++j) x[j] = f[j];
                    o.y[row] = refine ? 1.f : 0.f;
                }
            }
        }
        int next = 0;
        bool have = false;
        if (refine) {
            if (b.right < 0 && sp < kTileStack) S[sp++] = -(b.right + 1);
            if (b.left < 0) {
                next = -(b.left + 1);
                have = true;
            }
        }
        if (!have) {
            if (sp == 0) break;
            next = S[--sp];
        }
        k = uniform(next);
    }
}

// The learned traversal: per tile, up to 16 nodes popped together are
// decided by one MFMA pass of the network (all 64 lanes take part; lane
// i % 16 computes node i's features); a prediction p is taken when
// p >= conf (refine) or p <= 1 - conf (keep), and the error-bound criterion
// over the tile's pixels decides the rest.  Kept nodes and leaves add their
// estimate to every pixel.
__global__ __launch_bounds__(kCutThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_cut_learned_tile(CutEB t, const float4* __restrict__ gpos,
                                                                  const float4* __restrict__ gnrm,
                                                                  const float4* __restrict__ galb,
                                                                  const arc_material* __restrict__ mats, int W,
                                                                  int H, Threshold th, float eps, float delta,
                                                                  int budget, float conf,
                                                                  const float* __restrict__ params,
                                                                  const net::bf4* __restrict__ frag,
                                                                  float4* __restrict__ out_rad,
                                                                  unsigned long long* __restrict__ counters) {
    __shared__ int stk[kCutThreads / 64][kLearnStack];
    __shared__ float sb[3 * net::kHid + 1];
    __shared__ float snorm[2 * net::kIn];
    __shared__ net::bf4 sfrag[net::kF2b * 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int j = tid; j < net::kF2b * 64; j += kCutThreads) sfrag[j] = frag[j];
    for (int j = tid; j < net::kHid; j += kCutThreads) {
        sb[j] = params[net::kB1 + j];
        sb[net::kHid + j] = params[net::kB2 + j];
        sb[2 * net::kHid + j] = params[net::kW3 + j];
    }
    if (tid < 2 * net::kIn) snorm[tid] = params[net::kMean + tid];
    if (tid == 0) sb[3 * net::kHid] = params[net::kB3];
    __syncthreads();
    const int2 xy = tile_pixel(wave, lane, (W + 15) / 16, (H + 15) / 16);
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    const float thr = live ? th.at(q) : INFINITY;
    const TileStats T = tile_stats(P, live, thr);
    const int i = lane & 15, g = lane >> 4;
    const Corner K = tile_corner(P, thr, g);
    const int n = t.n;
    int* S = stk[wave];
    Acc acc;
    unsigned nvis = 0, nev = 0, nerr = 0, nbud = 0, nfb = 0;
    int sp = 0;
    if (T.live_frac > 0.f) S[sp++] = n >= 2 ? -1 : 0;
    while (sp > 0) {
        // keeps the network's LDS reads inside the loop: held in VGPRs across
        // it they would halve the occupancy
        asm volatile("" ::: "memory");
        const int cnt = min(kLearnBatch, sp);
        sp -= cnt;
        const int mine = S[sp + min(i, cnt - 1)];
        const bool have = i < cnt;
        const bool internal = have && mine < 0;
        float f[net::kIn];
        if (n >= 2) {  // uniform: every lane takes part (lanes without an internal node use the root)
            const int k = internal ? -(mine + 1) : 0;
            const float4* r = t.crec + kCrec * (size_t)(n + k);
            cut_features(T, K, t.nodes[k], r[0], r[1], r[2], r[3], eps, delta, f);
        } else {
#pragma unroll
            for (int j = 0; j < net::kIn; ++j) f[j] = 0.f;
        }
        net::bf4 x0, x1;
        net::input_frags_n(f, snorm, g, x0, x1);
        const float z = net::forward_lds(sfrag, sb, x0, x1, lane, g);
        const float pr = 1.f / (1.f + __expf(-z));
        const bool head = lane < kLearnBatch;
        uint64_t R = __ballot(head && internal && pr >= conf);
        uint64_t A = __ballot(head && have && (!internal || pr <= 1.f - conf));
        uint64_t U = __ballot(head && internal && pr < conf && pr > 1.f - conf);
        while (U) {  // not confident: the error-bound criterion over the tile's pixels
            const int j = __builtin_ctzll(U);
            U &= U - 1;
            const int k = -(__builtin_amdgcn_readlane(mine, j) + 1);
            const float4* r = t.crec + kCrec * (size_t)(n + k);
            const float crit = node_error(P, t.nodes[k], r[0], r[2], r[3], delta);
            if (__ballot(crit > thr)) R |= 1ull << j;
            else A |= 1ull << j;
            ++nfb;
        }
        if (R && (int)nev >= budget) {  // cut budget spent: keep the clusters
            nbud += __popcll(R);
            A |= R;
            R = 0;
        }
        while (R && sp + 2 * __popcll(R) > kLearnStack) {  // stack full: keep the last ones
            const uint64_t top = 1ull << (63 - __builtin_clzll(R));
            R &= ~top;
            A |= top;
            ++nerr;
        }
        if (head && ((R >> lane) & 1)) {
            const arc_bvh_node& b = t.nodes[-(mine + 1)];
            const int pos = sp + 2 * __popcll(R & ((1ull << lane) - 1));
            S[pos] = b.right;
            S[pos + 1] = b.left;
        }
        sp += 2 * __popcll(R);
        nvis += __popcll(R);
        while (A) {
            const int j = __builtin_ctzll(A);
            A &= A - 1;
            const int node = __builtin_amdgcn_readlane(mine, j);
            const int slot = node >= 0 ? node : n + (-(node + 1));
            const float4* r = t.crec + kCrec * (size_t)slot;
            acc.add_run(r[2], geom_est(P, r[0], r[1]));
            ++nev;
        }
        acc.maybe_flush();  // a batch keeps at most 16 nodes
    }
    acc.flush();
    const int alive = __popcll(__ballot(live));
    if (counters && lane == 0) {
        atomicAdd(&counters[C_VISITED], (unsigned long long)nvis * alive);
        atomicAdd(&counters[C_EVALUATED], (unsigned long long)nev * alive);
        atomicAdd(&counters[C_ERRORS], (unsigned long long)nerr);
        atomicAdd(&counters[C_BUDGET], (unsigned long long)nbud);
        atomicAdd(&counters[C_FALLBACK], (unsigned long long)nfb);
        atomicAdd(&counters[C_TESTED], (unsigned long long)nfb * alive);
    }
    if (live)
        out_rad[q] = make_float4(acc.k[0].s * P.br, acc.k[1].s * P.bg, acc.k[2].s * P.bb,
                                 counters ? (float)nev : 1.f);
}

}  // namespace

namespace arc {

namespace cut {

// Per-pixel estimates of the depth-d0 cut and the thresholds (the node
// records come with the tree, bvh_build_cut).
int prepare_cut(arc_gather* g, arc_bvh* b, const CutEBParams& cp, CutEB& t, Threshold& th) {
    arc_ctx* ctx = g->ctx;
    const int n = b->n;
    const int npix = g->width * g->height;
    if (n > (1 << 25)) return fail(ctx, ARC_E_UNSUPPORTED, "error-bound cut: more than 2^25 AVPLs");
    if (!b->has_cut_records || !g->crec.p || g->crec.bytes < sizeof(float4) * kCrec * (size_t)(2 * n - 1))
        return fail(ctx, ARC_E_STATE, "light cut: LBVH built without node records");
    ARC_TRY(ensure(ctx, g->cest, sizeof(float) * ((size_t)npix + 64)));
    float* est = g->cest.as<float>();
    float* sum = est + ((size_t)npix + 63) / 64 * 64;
    t = CutEB{b->nodes.as<arc_bvh_node>(), g->crec.as<float4>(), n};
    const dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    hipLaunchKernelGGL(k_cut_zero, dim3(1), dim3(64), 0, ctx->stream, sum);
    if (cp.d0 <= kEstListDepth) {
        const int bound = std::min(1 << cp.d0, n);  // the cut's nodes partition the n leaves
        hipLaunchKernelGGL(k_cut_estimate_list, grid, dim3(kCutThreads), sizeof(float4) * 2 * bound, ctx->stream,
                           t, cp.d0, g->gpos, g->gnrm, g->galb, g->mats.as<arc_material>(), g->width, g->height,
                           est, sum);
    } else {
        hipLaunchKernelGGL(k_cut_estimate, grid, dim3(kCutThreads), 0, ctx->stream, t, g->gpos, g->gnrm, g->galb,
                           g->mats.as<arc_material>(), g->width, g->height, cp.d0, est, sum);
    }
    th = Threshold{est, sum, cp.eps, cp.image_floor, 1.0f / (float)npix};
    return ARC_OK;
}

}  // namespace cut

int cut_error_bound(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                    hipEvent_t mark) {
    arc_ctx* ctx = g->ctx;
    CutEB t{};
    Threshold th{};
    ARC_TRY(prepare_cut(g, b, cp, t, th));
    const dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    if (mark) ARC_HIP(ctx, hipEventRecord(mark, ctx->stream));
    if (cp.per_tile) {
        // one instantiation per evaluation depth: the deeper sums would cost
        // the shallow ones registers (occupancy)
        auto tile = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(kCutThreads), 0, ctx->stream, t, g->gpos, g->gnrm, g->galb,
                               g->mats.as<arc_material>(), g->width, g->height, th, cp.delta, cp.budget,
                               cp.criterion, rad, counters);
        };
        switch (cp.eval_depth) {
            case 1: tile(k_cut_eb_tile<1>); break;
            case 2: tile(k_cut_eb_tile<2>); break;
            case 3: tile(k_cut_eb_tile<3>); break;
            case 4: tile(k_cut_eb_tile<4>); break;
            default: tile(k_cut_eb_tile<0>); break;
        }
    } else
        hipLaunchKernelGGL(k_cut_eb_px, grid, dim3(kCutThreads), 0, ctx->stream, t, g->gpos, g->gnrm, g->galb,
                           g->mats.as<arc_material>(), g->width, g->height, th, cp.delta, cp.budget, cp.criterion,
                           cp.eval_depth, rad, counters);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int cut_learned(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                hipEvent_t mark) {
    arc_ctx* ctx = g->ctx;
    if (!g->net || !g->net->ready) return fail(ctx, ARC_E_STATE, "learned cut without a trained arc_cutnet");
    CutEB t{};
    Threshold th{};
    ARC_TRY(prepare_cut(g, b, cp, t, th));
    const dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    if (mark) ARC_HIP(ctx, hipEventRecord(mark, ctx->stream));
    hipLaunchKernelGGL(k_cut_learned_tile, grid, dim3(kCutThreads), 0, ctx->stream, t, g->gpos, g->gnrm, g->galb,
                       g->mats.as<arc_material>(), g->width, g->height, th, cp.eps, cp.delta, cp.budget,
                       cp.confidence, g->net->params.as<float>(), g->net->frag.as<net::bf4>(), rad, counters);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int cut_samples(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float* X, float* y, int capacity, uint64_t seed,
                float keep, int* count) {
    arc_ctx* ctx = g->ctx;
    CutEB t{};
    Threshold th{};
    ARC_TRY(prepare_cut(g, b, cp, t, th));
    const dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    const double kd = std::min(1.0, (double)keep) * 4294967296.0;
    const LabelOut o{X, y, count, capacity, seed, kd >= 4294967295.0 ? 0xffffffffu : (uint32_t)kd};
    hipLaunchKernelGGL(k_cut_label_tile, grid, dim3(kCutThreads), 0, ctx->stream, t, b->ranges.as<int>(), g->gpos,
                       g->gnrm, g->galb, g->mats.as<arc_material>(), g->width, g->height, th, cp.eps, cp.delta, o);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

}  // namespace arc
