// Clustered antiradiance on gfx950 — the reference's clustered-deferred path
// (VisiblePointsBvh, bvh.cu:963-1208; kernel_gather_cluster_lights,
// cudaGather.cu:686-748; kernel_antiradiance_clusterred, :605-659).
//
// The reference clusters visible points, lists for every cluster the AVPLs
// whose incident ray hits the cluster's box grown by 2 photonRadius (ray-box
// slab test, cudaGather.cu:664-684), and sums each pixel over its cluster's
// list.  A disk hit needs the ray to pass within the photon radius of the
// pixel, so the list holds every AVPL that can contribute: the result equals
// the exact gather.  Here:
//   k_flag_count / k_scan_blocks / k_compact
//                  stable compaction of the AVPLs with non-zero antiradiance
//                  (the rest contribute nothing), no host round trip;
//   k_cluster_bounds
//                  one cluster per 32x32 pixel block, box of its G-buffer
//                  positions (the reference adds depth slices; a block of
//                  the screen with its tight box is what the list test needs);
//   k_cluster_tree
//                  a screen-space quadtree over the cluster boxes (2x2
//                  merges up to one root), one workgroup;
//   k_cluster_lists_tree
//                  lane = antiradiance AVPL, depth-first descent of the
//                  quadtree with a conservative test per box: ray-box slab
//                  test against the box grown by the photon radius plus a
//                  margin and the "pixel behind the AVPL's surface" bound
//                  (disk), or the cone against the box's bounding sphere;
//                  a ray meets a few of ~2000 clusters (a grazing one
//                  hundreds), so the descent tests a few dozen boxes where a
//                  flat loop tests all of them; appends are buffered per
//                  block in LDS and flushed by all of the block's threads.
//                  The reference's C x N list buffer becomes C x kCap with an
//                  overflow flag;
//   k_anti_clustered
//                  four 256-thread blocks per cluster (its 32x8-pixel rows),
//                  one wave per 8x8 tile; a block reads the cluster's list
//                  once and sorts its
//                  entries into per-tile sub-lists in LDS (ray against the
//                  tile's own box), then each wave walks its sub-list with
//                  wave-uniform (scalar) loads of the packed records and
//                  evaluates the guarded pair test per pixel; the list is
//                  filed and walked in chunks of the sub-lists' size, so no
//                  sub-list overflows.  A cluster whose list overflowed
//                  files every antiradiance AVPL the same way; the
//                  summation order within a pixel follows the list order,
//                  which the appends leave unspecified.
#include <cmath>

#include "gather_common.hpp"

using arc::f3;
using namespace arc::gm;
using namespace arc::gd;

namespace {

constexpr int kClusterPx = 32;   // cluster side in pixels
constexpr int kCap = 16384;      // list entries per cluster
constexpr int kCompact = 1024;   // compaction block
constexpr int kTiles = (kClusterPx / 8) * (kClusterPx / 8);  // 8x8 tiles per cluster
constexpr int kParts = 4;                                      // blocks per cluster (rows of 32 x 8 px)
constexpr int kBlockTiles = kTiles / kParts;                   // tiles per block
constexpr int kAntiThreads = 64 * kBlockTiles;                 // one wave per tile
constexpr int kSub = 1536;       // tile sub-list entries in LDS: list entries filed per chunk (default;
                                 // 768 and 3072 also built)
constexpr int kTreeMaxLevels = 10;  // cluster quadtree levels: up to 512 x 512 clusters

__device__ inline bool has_anti(const arc::AvplsDev& a, int i) {
    if (!a.anti) return false;
    const f3 A = a.anti[i];
    return A.x != 0.f || A.y != 0.f || A.z != 0.f;
}

__device__ inline uint32_t block_count(bool flag, uint32_t* sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    if (lane == 0) sh[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    return t;
}

__global__ __launch_bounds__(kCompact) void k_flag_count(arc::AvplsDev a, int* __restrict__ counts) {
    __shared__ uint32_t sh[kCompact / 64];
    const int i = blockIdx.x * kCompact + threadIdx.x;
    const bool f = i < a.n && has_anti(a, i);
    const uint32_t t = block_count(f, sh);
    if (threadIdx.x == 0) counts[blockIdx.x] = (int)t;
}

// Exclusive scan of nb block counts in one block; counts[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_blocks(int* __restrict__ counts, int nb) {
    __shared__ int sh[1024];
    int run = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int t = base + (int)threadIdx.x;
        const int v = t < nb ? counts[t] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int u = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += u;
            __syncthreads();
        }
        if (t < nb) counts[t] = run + sh[threadIdx.x] - v;
        run += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[nb] = run;
}

__global__ __launch_bounds__(kCompact) void k_compact(arc::AvplsDev a, const int* __restrict__ offs,
                                                      int* __restrict__ idx) {
    __shared__ uint32_t sh[kCompact / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * kCompact + threadIdx.x;
    const bool f = i < a.n && has_anti(a, i);
    const uint64_t m = __ballot(f);
    if (lane == 0) sh[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += sh[w];
    if (f) idx[offs[blockIdx.x] + before + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

// Box of the G-buffer positions of each 32x32 block: [C][8] = min xyz, -, max xyz, -.
__global__ __launch_bounds__(256) void k_cluster_bounds(const float4* __restrict__ gpos, int W, int H, int cw,
                                                        float* __restrict__ cbox) {
    __shared__ float s[6][256];
    const int c = blockIdx.x;
    const int x0 = (c % cw) * kClusterPx, y0 = (c / cw) * kClusterPx;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int k = threadIdx.x; k < kClusterPx * kClusterPx; k += 256) {
        const int x = x0 + (k % kClusterPx), y = y0 + (k / kClusterPx);
        if (x < W && y < H) {
            const float4 p = gpos[y * W + x];
            v[0] = fminf(v[0], p.x); v[1] = fminf(v[1], p.y); v[2] = fminf(v[2], p.z);
            v[3] = fmaxf(v[3], p.x); v[4] = fmaxf(v[4], p.y); v[5] = fmaxf(v[5], p.z);
        }
    }
    for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + off])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        cbox[c * 8 + k] = k == 3 || k == 7 ? 0.f : s[k < 3 ? k : k - 1][0];
    }
}

__global__ __launch_bounds__(256) void k_zero(int* __restrict__ p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

__device__ inline float safe_inv(float w) {
    // keeps the slab parameters finite (no 0 * inf) for axis-parallel rays
    const float m = fmaxf(fabsf(w), 1e-20f);
    return 1.0f / copysignf(m, w);
}

// Whether AVPL ray (p, w) with surface normal na can reach a pixel inside box
// b = (min xyz, max xyz).  Disk model: the forward ray must hit the box grown
// by the photon radius (a disk hit lies on the ray within that radius of a
// pixel), and part of the box must lie behind the AVPL's surface.  Cone
// model: the cone of half-angle theta about w must reach the box's bounding
// sphere: angle(c - a, w) <= theta + asin(R / |c - a|), with theta slightly
// enlarged (cos_t, sin_t).  Used for cluster boxes and for tile boxes: any
// box that contains the pixels' positions keeps every accepted pair.
template <int MODEL>
__device__ inline bool ray_box_hit(const f3& p, const f3& w, const f3& na, float ix, float iy, float iz,
                                   const float* b, float grow, float cos_t, float sin_t) {
    const float bx0 = b[0] - grow, by0 = b[1] - grow, bz0 = b[2] - grow;
    const float bx1 = b[3] + grow, by1 = b[4] + grow, bz1 = b[5] + grow;
    if (MODEL == ARC_ANTIRAD_DISK) {
        // ray-box slab test on the forward ray (intersectBB, cudaGather.cu:664-684)
        const float tx0 = (bx0 - p.x) * ix, tx1 = (bx1 - p.x) * ix;
        const float ty0 = (by0 - p.y) * iy, ty1 = (by1 - p.y) * iy;
        const float tz0 = (bz0 - p.z) * iz, tz1 = (bz1 - p.z) * iz;
        const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
        const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        // some point of the box lies behind the AVPL's surface: min n_a . (x - a) < 0
        const float behind = fminf(na.x * (bx0 - p.x), na.x * (bx1 - p.x)) +
                             fminf(na.y * (by0 - p.y), na.y * (by1 - p.y)) +
                             fminf(na.z * (bz0 - p.z), na.z * (bz1 - p.z));
        return tmax >= fmaxf(tmin, 0.f) && behind < grow;
    } else {
        const float cx = 0.5f * (b[0] + b[3]) - p.x, cy = 0.5f * (b[1] + b[4]) - p.y,
                    cz = 0.5f * (b[2] + b[5]) - p.z;
        const float ex = 0.5f * (b[3] - b[0]), ey = 0.5f * (b[4] - b[1]), ez = 0.5f * (b[5] - b[2]);
        const float R = sqrtf(ex * ex + ey * ey + ez * ez) + grow;
        const float D = sqrtf(cx * cx + cy * cy + cz * cz);
        const float wl = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
        const float cosang = (cx * w.x + cy * w.y + cz * w.z) / fmaxf(D * wl, 1e-30f);
        const float sb = fminf(R / fmaxf(D, 1e-30f), 1.f), cb = sqrtf(fmaxf(0.f, 1.f - sb * sb));
        return D <= R || cosang >= cos_t * cb - sin_t * sb - 1e-6f;
    }
}

// Quadtree over the cluster grid: level 0 is the clusters (cw x ch boxes,
// the k_cluster_bounds output), level l + 1 merges 2x2 nodes of level l, up
// to one root.  Boxes are [8] = min xyz, -, max xyz, -.
struct ClusterTree {
    int levels;
    int off[kTreeMaxLevels];  // first node of each level
    int w[kTreeMaxLevels], h[kTreeMaxLevels];
};

__global__ __launch_bounds__(1024) void k_cluster_tree(float* __restrict__ tree, ClusterTree t) {
    for (int l = 1; l < t.levels; ++l) {
        const int W = t.w[l], H = t.h[l], cw = t.w[l - 1], ch = t.h[l - 1];
        for (int k = threadIdx.x; k < W * H; k += blockDim.x) {
            const int x = k % W, y = k / W;
            float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            for (int cy = 2 * y; cy <= min(2 * y + 1, ch - 1); ++cy)
                for (int cx = 2 * x; cx <= min(2 * x + 1, cw - 1); ++cx) {
                    const float* c = tree + 8 * (size_t)(t.off[l - 1] + cy * cw + cx);
                    for (int e = 0; e < 3; ++e) {
                        b[e] = fminf(b[e], c[e]);
                        b[3 + e] = fmaxf(b[3 + e], c[4 + e]);
                    }
                }
            float* o = tree + 8 * (size_t)(t.off[l] + k);
            o[0] = b[0]; o[1] = b[1]; o[2] = b[2]; o[3] = 0.f;
            o[4] = b[3]; o[5] = b[4]; o[6] = b[5]; o[7] = 0.f;
        }
        __syncthreads();  // level l complete before level l + 1 reads it
    }
}

// Lists by tree descent: lane = antiradiance AVPL, depth-first through the
// quadtree with the same conservative test at every level (a parent box
// contains its children, so a rejected parent rejects every cluster below).
// A popped node's children are loaded together and tested together, so a
// lane waits on one round of loads per node and not one per child.  A ray
// that sweeps the screen reaches hundreds of clusters; its appends go to a
// block-wide LDS buffer (cluster << 8 | lane) instead of one global atomic
// round trip each, and once every lane of the block has finished its
// descent the block's threads share the global appends.  Entries past the
// buffer's end are appended directly.  The descent stack (one column per
// thread, 1 + 3 (levels - 1) entries) is dynamic LDS.
constexpr int kListThreads = 256;
constexpr int kPairBuf = 4096;  // buffered appends per block
static_assert(kListThreads == 256, "an append packs the lane in 8 bits");

template <int MODEL>
__global__ __launch_bounds__(kListThreads) void k_cluster_lists_tree(arc::AvplsDev a, const int* __restrict__ idx,
                                                                     const int* __restrict__ n_anti_ptr,
                                                                     const float4* __restrict__ tree, ClusterTree t,
                                                                     float grow, float cos_t, float sin_t,
                                                                     int* __restrict__ counts, int* __restrict__ lists) {
    __shared__ int s_off[kTreeMaxLevels], s_w[kTreeMaxLevels], s_h[kTreeMaxLevels];
    __shared__ int s_pair[kPairBuf];
    __shared__ int s_n;
    extern __shared__ int stk[];  // [depth][kListThreads]
    const int tid = threadIdx.x;
    if (tid < kTreeMaxLevels) {
        s_off[tid] = t.off[tid];
        s_w[tid] = t.w[tid];
        s_h[tid] = t.h[tid];
    }
    if (tid == 0) s_n = 0;
    __syncthreads();
    const int n_anti = *n_anti_ptr;
    const int base = blockIdx.x * kListThreads;
    const int j = base + tid;
    auto append = [&](int c, int i) {
        const int k = atomicAdd(&s_n, 1);
        if (k < kPairBuf) {
            s_pair[k] = (c << 8) | tid;
        } else {
            const int slot = atomicAdd(&counts[c], 1);
            if (slot < kCap) lists[(size_t)c * kCap + slot] = i;
        }
    };
    if (j < n_anti) {
        const int i = idx[j];
        const f3 p = a.pos[i], w = a.dir[i], na = a.nrm[i];
        const float ix = safe_inv(w.x), iy = safe_inv(w.y), iz = safe_inv(w.z);
        const int top = t.levels - 1;
        const float4 r0 = tree[2 * s_off[top]], r1 = tree[2 * s_off[top] + 1];
        const float rb[6] = {r0.x, r0.y, r0.z, r1.x, r1.y, r1.z};
        if (ray_box_hit<MODEL>(p, w, na, ix, iy, iz, rb, grow, cos_t, sin_t)) {
            if (top == 0) {
                append(0, i);
            } else {
                int sp = 0;
                stk[(sp++) * kListThreads + tid] = top << 24;  // (level, y, x) of a node whose box was hit
                while (sp > 0) {
                    const int v = stk[(--sp) * kListThreads + tid];
                    const int l = (v >> 24) - 1, y = 2 * ((v >> 12) & 4095), x = 2 * (v & 4095);
                    const int cw = s_w[l], ch = s_h[l], off = s_off[l];
                    const bool ok[4] = {true, x + 1 < cw, y + 1 < ch, x + 1 < cw && y + 1 < ch};
                    float4 lo[4], hi[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {  // a missing child loads the first one's box
                        const int k = off + (ok[q] ? (y + (q >> 1)) * cw + x + (q & 1) : y * cw + x);
                        lo[q] = tree[2 * k];
                        hi[q] = tree[2 * k + 1];
                    }
#pragma unroll
                    for (int q = 3; q >= 0; --q) {  // pushed in reverse: popped in row order
                        const float b[6] = {lo[q].x, lo[q].y, lo[q].z, hi[q].x, hi[q].y, hi[q].z};
                        if (!ok[q] || !ray_box_hit<MODEL>(p, w, na, ix, iy, iz, b, grow, cos_t, sin_t)) continue;
                        const int cy = y + (q >> 1), cx = x + (q & 1);
                        if (l == 0) append(cy * cw + cx, i);
                        else stk[(sp++) * kListThreads + tid] = (l << 24) | (cy << 12) | cx;
                    }
                }
            }
        }
    }
    __syncthreads();
    const int np = min(s_n, kPairBuf);
    for (int k = tid; k < np; k += kListThreads) {
        const int e = s_pair[k], c = e >> 8;
        const int slot = atomicAdd(&counts[c], 1);
        if (slot < kCap) lists[(size_t)c * kCap + slot] = idx[base + (e & 255)];
    }
}

// Four blocks per 32x32 cluster (its rows of 32x8 pixels), one wave per 8x8
// tile.  A block reads the cluster's list once: each thread takes an entry
// and tests its ray against the block's 4 tile boxes (the same conservative
// test as the lists, on smaller boxes), appending hits to the tile's
// sub-list in LDS.  Each wave then walks its tile's sub-list with
// wave-uniform (scalar) loads of the packed records.
// LDS at 25 KB per block, so the occupancy is set by the registers (at least 5 waves
// per SIMD) and the blocks share the CUs with the cut kernel that runs
// beside them on the other stream.
template <int MODEL, bool GENERAL, int kChunk>
__global__ __launch_bounds__(kAntiThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_anti_clustered(
    const float4* __restrict__ rec, arc::AvplsDev raw, const float4* __restrict__ gpos,
    const float4* __restrict__ gnrm, const float4* __restrict__ galb, const arc_material* __restrict__ mats, int W,
    int H, int cw, float camx, float camy, float camz, AntiParams ap, const int* __restrict__ counts,
    const int* __restrict__ lists, const int* __restrict__ idx, const int* __restrict__ n_anti_ptr, float grow,
    float cos_t, float sin_t, float* __restrict__ part, unsigned long long* __restrict__ counters) {
    __shared__ float tbox[kBlockTiles][6];
    __shared__ float tface[kBlockTiles][8];  // cone: m xyz, h, s, box centre xyz
    __shared__ int tlive[kBlockTiles];
    __shared__ int tcount[kBlockTiles];
    __shared__ int sub[kBlockTiles][kChunk];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int c = b / kParts, half = b % kParts;
    const int x = (c % cw) * kClusterPx + (wave & 3) * 8 + (lane & 7);
    const int y = (c / cw) * kClusterPx + half * (kClusterPx / kParts) + (wave >> 2) * 8 + (lane >> 3);
    const bool live = x < W && y < H;
    const int npix = W * H;
    const int q = live ? y * W + x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    {  // the tile's box
        const float big = INFINITY;
        float tb[6] = {live ? P.px : big, live ? P.py : big, live ? P.pz : big,
                       live ? P.px : -big, live ? P.py : -big, live ? P.pz : -big};
        for (int off = 32; off > 0; off >>= 1) {
            for (int k = 0; k < 3; ++k) tb[k] = fminf(tb[k], __shfl_xor(tb[k], off));
            for (int k = 3; k < 6; ++k) tb[k] = fmaxf(tb[k], __shfl_xor(tb[k], off));
        }
        const bool any = __ballot(live) != 0;
        if (lane == 0) {
            for (int k = 0; k < 6; ++k) tbox[wave][k] = tb[k];
            tlive[wave] = any ? 1 : 0;  // a tile wholly outside the image has nothing to sum
            tcount[wave] = 0;
        }
        if (MODEL == ARC_ANTIRAD_CONE) {
            // the tile's normal cone: unit axis m (the normals' mean
            // direction), s = max |n - m| and h = min m . p over its live
            // pixels; a cone pair needs n . (a - p) > 0, and
            // n . (a - p) <= m . a - h + s (|a - c| + R) for every pixel
            float mx = live ? P.nx : 0.f, my = live ? P.ny : 0.f, mz = live ? P.nz : 0.f;
            for (int off = 32; off > 0; off >>= 1) {
                mx += __shfl_xor(mx, off);
                my += __shfl_xor(my, off);
                mz += __shfl_xor(mz, off);
            }
            const float ml = sqrtf(mx * mx + my * my + mz * mz);
            if (ml > 1e-20f) {
                mx /= ml; my /= ml; mz /= ml;
            } else {
                mx = 0.f; my = 0.f; mz = 1.f;  // any unit axis: s then covers every normal
            }
            const float ex = P.nx - mx, ey = P.ny - my, ez = P.nz - mz;
            float s2 = live ? ex * ex + ey * ey + ez * ez : 0.f;
            float h = live ? mx * P.px + my * P.py + mz * P.pz : INFINITY;
            for (int off = 32; off > 0; off >>= 1) {
                s2 = fmaxf(s2, __shfl_xor(s2, off));
                h = fminf(h, __shfl_xor(h, off));
            }
            if (lane == 0) {
                tface[wave][0] = mx; tface[wave][1] = my; tface[wave][2] = mz;
                tface[wave][3] = h;
                tface[wave][4] = sqrtf(s2) * 1.0001f + 1e-6f;
                tface[wave][5] = 0.5f * (tb[0] + tb[3]);
                tface[wave][6] = 0.5f * (tb[1] + tb[4]);
                tface[wave][7] = 0.5f * (tb[2] + tb[5]);
            }
        }
    }
    __syncthreads();
    const int listed = counts[c];
    const bool overflow = listed > kCap;
    const int count = overflow ? *n_anti_ptr : listed;
    const int* list = overflow ? idx : lists + (size_t)c * kCap;
    // Entries are fetched 64 at a time, one per lane, and broadcast with
    // readlane.  Their records are read with scalar loads four entries at a
    // time: scalar loads complete out of order, so one wait covers a batch
    // and the four loads overlap.  Each chunk of 64 is one plain FP32 run,
    // Kahan-added.  (Staging each chunk's records in LDS instead measured
    // slower: 3.8 -> 49 ms for the disk model on the 1080p x 1M frame.)
    static_assert(kRun == 64, "one compensated add per 64-entry chunk");
    auto pair = [&](int i, const float4& r0, const float4& r1, const float4& r2v, const float4& r4, float& s0,
                    float& s1, float& s2) {
        const float dx = r0.x - P.px, dy = r0.y - P.py, dz = r0.z - P.pz;
        const float dd = dx * dx + dy * dy + dz * dz;
        // hardware reciprocal square root (about 1 ulp): it scales the guard
        // band and the cone's 1 / d^2, never a decision inside the band
        const float inv = __builtin_amdgcn_rsqf(dd);
        const float cn = P.nx * dx + P.ny * dy + P.nz * dz;
        const float ca = -(r1.x * dx + r1.y * dy + r1.z * dz);
        const float m = anti_weight<MODEL>(P, r0, r1, r2v, r4.w, dx, dy, dz, dd, inv, cn, ca, ap, gpos, gnrm, q, raw, i);
        if (GENERAL) {
            if (m != 0.f) {
                using V3 = V<float>;
                Mat<float> pm = material<float>(mats, P.mat);
                if (galb) pm.diffuse = mk(P.br / kOneOverPi, P.bg / kOneOverPi, P.bb / kOneOverPi);
                const V3 b = f_r3(mk(r0.x, r0.y, r0.z), mk(P.px, P.py, P.pz), mk(camx, camy, camz),
                                  mk(P.nx, P.ny, P.nz), pm) * m;
                s0 += r4.x * b.x; s1 += r4.y * b.y; s2 += r4.z * b.z;
            }
        } else {
            s0 += r4.x * m; s1 += r4.y * m; s2 += r4.z * m;
        }
    };
    // The cluster's list is taken kChunk entries at a time: the block files
    // each chunk into the tiles' sub-lists (a chunk cannot overflow them),
    // then each wave walks its tile's sub-list, so a tile never walks entries
    // its own box test rejected.
    Kahan ka[3];
    unsigned long long walked = 0;
    for (int cbase = 0; cbase < count; cbase += kChunk) {
        const int cend = min(count, cbase + kChunk);
        if (lane == 0) tcount[wave] = 0;
        __syncthreads();  // counts cleared; the previous chunk's sub-lists are read
        for (int j = cbase + tid; j < cend; j += kAntiThreads) {
            const int i = list[j];
            const f3 p = raw.pos[i], w = raw.dir[i], na = raw.nrm[i];
            const float ix = safe_inv(w.x), iy = safe_inv(w.y), iz = safe_inv(w.z);
            for (int t = 0; t < kBlockTiles; ++t) {
                if (MODEL == ARC_ANTIRAD_CONE && tlive[t]) {
                    // behind every pixel's tangent plane: no pixel of the tile faces the AVPL
                    const float* f = tface[t];
                    const float* bb = tbox[t];
                    const float qx = p.x - f[5], qy = p.y - f[6], qz = p.z - f[7];
                    const float bx = bb[3] - bb[0], by = bb[4] - bb[1], bz = bb[5] - bb[2];
                    const float reach = sqrtf(qx * qx + qy * qy + qz * qz) + 0.5f * sqrtf(bx * bx + by * by + bz * bz);
                    if (f[0] * p.x + f[1] * p.y + f[2] * p.z - f[3] + f[4] * reach < -grow) continue;
                }
                if (tlive[t] && ray_box_hit<MODEL>(p, w, na, ix, iy, iz, tbox[t], grow, cos_t, sin_t)) {
                    const int slot = atomicAdd(&tcount[t], 1);
                    sub[t][slot] = i;  // slot < cend - cbase <= kChunk
                }
            }
        }
        __syncthreads();
        const int walk = tlive[wave] ? tcount[wave] : 0;
        walked += walk;
        for (int base = 0; base < walk; base += 64) {
            const int jj = base + lane;
            // lanes past the end hold the chunk's first entry (a valid index)
            const int mine = jj < walk ? sub[wave][jj] : sub[wave][base];
            const int cnt = __builtin_amdgcn_readfirstlane(min(64, walk - base));
            float s0 = 0.f, s1 = 0.f, s2 = 0.f;
            for (int u = 0; u < cnt; u += 4) {
                int i[4];
                float4 a0[4], a1[4], a2[4], a4[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    i[k] = __builtin_amdgcn_readlane(mine, min(u + k, cnt - 1));
                    const float4* r = rec + (size_t)kRecF4 * i[k];
                    a0[k] = r[0];
                    a1[k] = r[1];
                    a2[k] = r[2];
                    a4[k] = r[4];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k == 0 || u + k < cnt) pair(i[k], a0[k], a1[k], a2[k], a4[k], s0, s1, s2);
            }
            ka[0].add(s0); ka[1].add(s1); ka[2].add(s2);
        }
    }
    const int alive = __popcll(__ballot(live));  // all lanes take part in the ballot
    if (counters && lane == 0) {
        if (wave == 0) atomicAdd(&counters[C_ANTI_VISITED], (unsigned long long)count * kBlockTiles);
        atomicAdd(&counters[C_ANTI_EVALUATED], walked * alive);
    }
    if (!live) return;
    float* o = part + q;
    if (GENERAL) {
        o[3 * (size_t)npix] = ka[0].s; o[4 * (size_t)npix] = ka[1].s; o[5 * (size_t)npix] = ka[2].s;
    } else {
        o[3 * (size_t)npix] = ka[0].s * P.br; o[4 * (size_t)npix] = ka[1].s * P.bg;
        o[5 * (size_t)npix] = ka[2].s * P.bb;
    }
}

template <int MODEL, bool GENERAL>
void launch_pixels(arc_gather* g, const arc::AvplsDev& a, int cw, int C, float3 cam, const AntiParams& ap,
                   const int* n_anti_ptr, float grow, float cos_t, float sin_t, float* part,
                   unsigned long long* counters) {
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(C * kParts), dim3(kAntiThreads), 0, g->ctx->stream, g->rec.as<float4>(), a,
                           g->gpos, g->gnrm, g->galb, g->mats.as<arc_material>(), g->width, g->height, cw, cam.x,
                           cam.y, cam.z, ap, g->ccount.as<int>(), g->clist.as<int>(), g->anti_idx.as<int>(),
                           n_anti_ptr, grow, cos_t, sin_t, part, counters);
    };
    // list entries filed per pass: a launch-shape setting (LDS against passes)
    if (g->anti_chunk == 768) launch(k_anti_clustered<MODEL, GENERAL, 768>);
    else if (g->anti_chunk == 3072) launch(k_anti_clustered<MODEL, GENERAL, 3072>);
    else launch(k_anti_clustered<MODEL, GENERAL, kSub>);
}

}  // namespace

namespace arc {

int anti_clustered(arc_gather* g, const AvplsDev& a, int model, float3 cam, const AntiParams& ap,
                   double half_angle, float extent, float* part, unsigned long long* counters, hipEvent_t mark) {
    arc_ctx* ctx = g->ctx;
    const int n = a.n;
    const int nb = (n + kCompact - 1) / kCompact;
    const int cw = (g->width + kClusterPx - 1) / kClusterPx, ch = (g->height + kClusterPx - 1) / kClusterPx;
    const int C = cw * ch;
    // anti_idx: [n] compacted indices, then [nb + 1] block offsets (last = count)
    ARC_TRY(ensure(ctx, g->anti_idx, sizeof(int) * ((size_t)n + nb + 1)));
    ClusterTree tree{};
    tree.w[0] = cw;
    tree.h[0] = ch;
    int nodes = C;
    tree.levels = 1;
    while (tree.w[tree.levels - 1] > 1 || tree.h[tree.levels - 1] > 1) {
        if (tree.levels == kTreeMaxLevels) return fail(ctx, ARC_E_UNSUPPORTED, "image too large for the cluster tree");
        const int l = tree.levels++;
        tree.off[l] = nodes;
        tree.w[l] = (tree.w[l - 1] + 1) / 2;
        tree.h[l] = (tree.h[l - 1] + 1) / 2;
        nodes += tree.w[l] * tree.h[l];
    }
    ARC_TRY(ensure(ctx, g->cbox, sizeof(float) * 8 * (size_t)nodes));
    ARC_TRY(ensure(ctx, g->ccount, sizeof(int) * (size_t)C));
    ARC_TRY(ensure(ctx, g->clist, sizeof(int) * (size_t)C * kCap));
    int* idx = g->anti_idx.as<int>();
    int* offs = idx + n;
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(kCompact), 0, ctx->stream, a, offs);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, ctx->stream, offs, nb);
    hipLaunchKernelGGL(k_compact, dim3(nb), dim3(kCompact), 0, ctx->stream, a, offs, idx);
    hipLaunchKernelGGL(k_cluster_bounds, dim3(C), dim3(256), 0, ctx->stream, g->gpos, g->width, g->height, cw,
                       g->cbox.as<float>());
    hipLaunchKernelGGL(k_zero, dim3((C + 255) / 256), dim3(256), 0, ctx->stream, g->ccount.as<int>(), C);
    ARC_LAUNCHED(ctx);
    // The list test must keep every pair the FP32 / FP64 pair test accepts.
    // For disks the box grows by the photon radius plus a relative and an
    // absolute margin (the line may pass the pixel that far away); the cone
    // test has no radius, so its box grows only by a rounding margin, 2e-5 of
    // the scene's extent (FP32 carries about 1e-7 of it).
    const float grow = model == ARC_ANTIRAD_DISK ? ap.radius * 1.001f + 1e-5f
                                                 : 2e-5f * fmaxf(extent, 1e-30f) + 1e-30f;
    const double th = std::min(M_PI, half_angle * 1.001 + 1e-4);
    const float cos_t = (float)std::cos(th), sin_t = (float)std::sin(th);
    const int* n_anti = offs + nb;
    if (tree.levels > 1)
        hipLaunchKernelGGL(k_cluster_tree, dim3(1), dim3(1024), 0, ctx->stream, g->cbox.as<float>(), tree);
    const dim3 lgrid((n + kListThreads - 1) / kListThreads);
    const size_t lstack = sizeof(int) * kListThreads * (size_t)(1 + 3 * (tree.levels - 1));
    const float4* tbox = reinterpret_cast<const float4*>(g->cbox.as<float>());
    if (model == ARC_ANTIRAD_DISK)
        hipLaunchKernelGGL(k_cluster_lists_tree<ARC_ANTIRAD_DISK>, lgrid, dim3(kListThreads), lstack, ctx->stream, a,
                           idx, n_anti, tbox, tree, grow, cos_t, sin_t, g->ccount.as<int>(), g->clist.as<int>());
    else
        hipLaunchKernelGGL(k_cluster_lists_tree<ARC_ANTIRAD_CONE>, lgrid, dim3(kListThreads), lstack, ctx->stream, a,
                           idx, n_anti, tbox, tree, grow, cos_t, sin_t, g->ccount.as<int>(), g->clist.as<int>());
    ARC_LAUNCHED(ctx);
    if (mark) ARC_HIP(ctx, hipEventRecord(mark, ctx->stream));
    if (model == ARC_ANTIRAD_DISK) {
        if (g->general) launch_pixels<ARC_ANTIRAD_DISK, true>(g, a, cw, C, cam, ap, n_anti, grow, cos_t, sin_t, part, counters);
        else launch_pixels<ARC_ANTIRAD_DISK, false>(g, a, cw, C, cam, ap, n_anti, grow, cos_t, sin_t, part, counters);
    } else {
        if (g->general) launch_pixels<ARC_ANTIRAD_CONE, true>(g, a, cw, C, cam, ap, n_anti, grow, cos_t, sin_t, part, counters);
        else launch_pixels<ARC_ANTIRAD_CONE, false>(g, a, cw, C, cam, ap, n_anti, grow, cos_t, sin_t, part, counters);
    }
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

}  // namespace arc
