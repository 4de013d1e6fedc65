// Karras LBVH on gfx950 — replaces Bvh::create and its AvplBvh / SimpleBvh
// specialisations (bvh.cu:325-454, 673-808, 859-907) and morton.h.
//
// Pipeline (one stream, no host round trip):
//   k_bounds_*   per-axis min / max of positions and normals, init FLT_MAX /
//                FLT_MIN as the reference's thrust::reduce (bvh.cu:346-361)
//   k_encode     (v - min) / (max - min) with IEEE division, 10-bit
//                quantisation, 30-bit interleave (morton.h:34-50, 82-92) or
//                the 60-bit 6D code truncated to 32 bits (morton.h:94-107,
//                bvh.h:17); value = input index
//   sort_pairs   stable radix sort, so equal codes keep input order
//   k_permute    sorted leaf positions / normals and leaf node data
//   k_karras     one thread per internal node (bvh.cu:188-278)
//   k_refit_*    bottom-up AABB merge, one thread per leaf, block-local
//                counters in LDS and a one-workgroup pass over the nodes that
//                span blocks; the AvplBvh variant also forms cluster
//                representatives (bvh.cu:726-744)
//
// This file is compiled with -ffp-contract=off: the refit's float results
// (cluster radiance sums, representative choice) must match the oracle bit
// for bit, so no multiply-add may be fused.
#include <cfloat>

#include "arc_internal.hpp"

namespace {

using arc::f3;
using arc::kCrec;
using arc::kMrec;
using arc::write_cut_record;

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 512;

__device__ inline f3 ldf3(const f3* p, size_t i) { return p[i]; }

// ------------------------------------------------------------ bounds ----
// partials: [block][12] = pos min xyz, pos max xyz, nrm min xyz, nrm max xyz
__global__ __launch_bounds__(kBlock) void k_bounds_partial(const f3* __restrict__ pos,
                                                           const f3* __restrict__ nrm, int n,
                                                           float* __restrict__ partials) {
    float v[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = FLT_MAX;
        v[3 + k] = FLT_MIN;  // the reference's max-init (Q1), a tiny positive value
        v[6 + k] = FLT_MAX;
        v[9 + k] = FLT_MIN;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const f3 p = pos[i];
        v[0] = fminf(v[0], p.x); v[1] = fminf(v[1], p.y); v[2] = fminf(v[2], p.z);
        v[3] = fmaxf(v[3], p.x); v[4] = fmaxf(v[4], p.y); v[5] = fmaxf(v[5], p.z);
        if (nrm) {
            const f3 q = nrm[i];
            v[6] = fminf(v[6], q.x); v[7] = fminf(v[7], q.y); v[8] = fminf(v[8], q.z);
            v[9] = fmaxf(v[9], q.x); v[10] = fmaxf(v[10], q.y); v[11] = fmaxf(v[11], q.z);
        }
    }
    __shared__ float s[12][kBlock];
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                const float a = s[k][threadIdx.x], b = s[k][threadIdx.x + off];
                s[k][threadIdx.x] = ((k / 3) & 1) ? fmaxf(a, b) : fminf(a, b);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) partials[blockIdx.x * 12 + threadIdx.x] = s[threadIdx.x][0];
}

// One wave per component: lanes stride over the partials, then a shuffle tree.
__global__ __launch_bounds__(768) void k_bounds_final(const float* __restrict__ partials, int nparts,
                                                      float* __restrict__ bounds) {
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool is_max = (k / 3) & 1;
    float r = is_max ? FLT_MIN : FLT_MAX;
    for (int b = lane; b < nparts; b += 64) {
        const float x = partials[b * 12 + k];
        r = is_max ? fmaxf(r, x) : fminf(r, x);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(r, off);
        r = is_max ? fmaxf(r, o) : fminf(r, o);
    }
    if (lane == 0) bounds[k] = r;
}

// ------------------------------------------------------------ Morton ----
__device__ inline uint32_t expand2(uint32_t x) {  // morton.h:34-42
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0xc30c30c3u;
    x = (x ^ (x << 2)) & 0x49249249u;
    return x;
}

__device__ inline uint32_t quant10(float c) {
    // fminf(fmaxf(c * 1024, 0), 1023), truncating conversion; NaN -> 0
    return (uint32_t)fminf(fmaxf(c * 1024.0f, 0.0f), 1023.0f);
}

// 6-way interleave of 10-bit inputs, bit i of input k at 6 i + k
// (morton.h:67-80); only the low 32 bits survive (MC_SIZE = uint32_t).
__device__ inline uint32_t interleave6_lo32(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                            uint32_t e, uint32_t f) {
    uint64_t r = 0;
    const uint32_t in[6] = {a, b, c, d, e, f};
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k) r |= (uint64_t)((in[k] >> i) & 1u) << (6 * i + k);
    return (uint32_t)r;
}

__global__ __launch_bounds__(kBlock) void k_encode(const f3* __restrict__ pos, const f3* __restrict__ nrm,
                                                   int n, const float* __restrict__ bounds,
                                                   int consider_normals, uint32_t* __restrict__ keys,
                                                   int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 p = pos[i];
    // Normalize functor (bvh.cu:73-81): component-wise IEEE division
    const float x = (p.x - bounds[0]) / (bounds[3] - bounds[0]);
    const float y = (p.y - bounds[1]) / (bounds[4] - bounds[1]);
    const float z = (p.z - bounds[2]) / (bounds[5] - bounds[2]);
    uint32_t code;
    if (consider_normals) {
        const f3 q = nrm[i];
        const float u = (q.x - bounds[6]) / (bounds[9] - bounds[6]);
        const float v = (q.y - bounds[7]) / (bounds[10] - bounds[7]);
        const float w = (q.z - bounds[8]) / (bounds[11] - bounds[8]);
        code = interleave6_lo32(quant10(x), quant10(y), quant10(z), quant10(u), quant10(v), quant10(w));
    } else {
        code = (expand2(quant10(z)) << 2) + (expand2(quant10(y)) << 1) + expand2(quant10(x));
    }
    keys[i] = code;
    vals[i] = i;
}

// Flags AVPLs with bounce <= 0 (they would misalign leaves and node data).
__global__ __launch_bounds__(kBlock) void k_check_bounce(const int* __restrict__ bounce, int n,
                                                         int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && bounce[i] <= 0) atomicOr(flag, 1);
}

// --------------------------------------------------------- permute ----
__global__ __launch_bounds__(kBlock) void k_permute_points(const int* __restrict__ perm,
                                                           const f3* __restrict__ pos,
                                                           const f3* __restrict__ nrm, int n,
                                                           f3* __restrict__ spos, f3* __restrict__ snrm) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int o = perm[s];
    spos[s] = pos[o];
    if (nrm) snrm[s] = nrm[o];
}

struct NodeData {  // AvplBvhNodeDataParam over 2n-1 slots
    int* size;
    int* mat;
    const float* rnd;
    f3* pos;
    f3* nrm;
    f3* rad;
    f3* dir;
};

// bvh_build_cut's per-slot records: merge sums (kMrec float4) and the cut's
// node records (kCrec float4).
struct CutRecs {
    float4* merge;
    float4* crec;
};

// The payload of the sort gathered by the permutation: read at random, four
// SoA arrays of 12-byte elements cost a cache line each per AVPL, so they
// are first packed in input order into one 64-byte record per AVPL
// (position and material, normal, radiance, direction).
__global__ __launch_bounds__(kBlock) void k_stage_avpls(arc::AvplsDev a, float4* __restrict__ staged) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const f3 p = a.pos[i], q = a.nrm[i], L = a.rad[i], w = a.dir[i];
    float4* r = staged + 4 * (size_t)i;
    r[0] = make_float4(p.x, p.y, p.z, __int_as_float(a.mat[i]));
    r[1] = make_float4(q.x, q.y, q.z, 0.f);
    r[2] = make_float4(L.x, L.y, L.z, 0.f);
    r[3] = make_float4(w.x, w.y, w.z, 0.f);
}

__global__ __launch_bounds__(kBlock) void k_permute_avpls(const int* __restrict__ perm, int n,
                                                          const float4* __restrict__ staged,
                                                          f3* __restrict__ spos, f3* __restrict__ snrm, NodeData nd) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4* rec = staged + 4 * (size_t)perm[s];
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const f3 p{r0.x, r0.y, r0.z}, q{r1.x, r1.y, r1.z}, L{r2.x, r2.y, r2.z}, w{r3.x, r3.y, r3.z};
    const int mat = __float_as_int(r0.w);
    spos[s] = p;
    snrm[s] = q;
    nd.size[s] = 1;
    nd.mat[s] = mat;
    nd.pos[s] = p;
    nd.nrm[s] = q;
    nd.rad[s] = L;
    nd.dir[s] = w;
}

// The leaves of bvh_build_cut: position (for the box refit) and normal, the
// merge sums and the node record of each sorted AVPL.  L x f_r of the AVPL's Lambertian
// material (f_r = 1 for a light-source AVPL, w = 0, cudaGather.cu:89-90); the
// table holds 256 entries.
__global__ __launch_bounds__(kBlock) void k_permute_cut(const int* __restrict__ perm, int n,
                                                        const float4* __restrict__ staged, f3* __restrict__ spos,
                                                        f3* __restrict__ snrm, const arc_material* __restrict__ mats,
                                                        CutRecs cr) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4* rec = staged + 4 * (size_t)perm[s];
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const f3 p{r0.x, r0.y, r0.z}, q{r1.x, r1.y, r1.z}, L{r2.x, r2.y, r2.z};
    spos[s] = p;
    snrm[s] = q;
    f3 r = L;
    if (r3.x != 0.f || r3.y != 0.f || r3.z != 0.f) {
        const arc_material& m = mats[__float_as_int(r0.w) & 255];
        r = f3{L.x * (0.318309886f * m.diffuse.x), L.y * (0.318309886f * m.diffuse.y),
               L.z * (0.318309886f * m.diffuse.z)};
    }
    const float w8 = (r.x + r.y + r.z) * (1.0f / 3.0f);
    const float ql = q.x * q.x + q.y * q.y + q.z * q.z;
    const float iq = ql > 0.f ? 1.0f / sqrtf(ql) : 0.f;
    const float4 mo = make_float4(w8 * p.x, w8 * p.y, w8 * p.z, w8);
    const f3 nm{w8 * q.x * iq, w8 * q.y * iq, w8 * q.z * iq};
    float4* mg = cr.merge + kMrec * (size_t)s;
    mg[0] = make_float4(r.x, r.y, r.z, __int_as_float(1));
    mg[1] = mo;
    mg[2] = make_float4(nm.x, nm.y, nm.z, 0.f);
    write_cut_record(cr.crec + kCrec * (size_t)s, r, mo, nm, 1, 0.f, p, q);
}

__global__ __launch_bounds__(kBlock) void k_random(uint64_t seed, int count, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = arc::hash_uniform(seed, (uint64_t)i);
}

// ---------------------------------------------------------- Karras ----
__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* __restrict__ m, int num_nodes,
                                                   arc_bvh_node* __restrict__ nodes,
                                                   int* __restrict__ parents, int* __restrict__ ranges,
                                                   int* __restrict__ nready) {
    // The searches of most nodes stay within 128 slots of the node (the first
    // l_max): the block's codes and 256 on each side are read once into LDS,
    // and a probe outside that window reads global memory.
    constexpr int kWin = 3 * kBlock;
    __shared__ uint32_t win[kWin];
    const int wbase = (int)(blockIdx.x * blockDim.x) - kBlock;
    for (int k = threadIdx.x; k < kWin; k += blockDim.x) {
        const int j = wbase + k;
        win[k] = (j >= 0 && j <= num_nodes) ? m[j] : 0u;
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) nready[0] = nready[1] = nready[2] = 0;  // the refits' list lengths
    if (i >= num_nodes) return;             // no barrier follows
    const uint32_t mi = win[i - wbase];
    // delta(i, j) (bvh.cu:124-135): -1 outside [0, num_nodes]; on equal codes
    // the common prefix continues into the slot indices
    auto dl = [&](int j) -> int {
        if (j < 0 || j > num_nodes) return -1;
        const int k = j - wbase;
        const uint32_t x = mi ^ ((k >= 0 && k < kWin) ? win[k] : m[j]);
        return x ? __clz(x) : 32 + __clz((uint32_t)(i ^ j));
    };
    const int root_slot = num_nodes + 1;
    // direction of the range: sign(0) = +1 (bvh.cu:170-173, 203-204)
    const int d = (dl(i + 1) - dl(i - 1)) < 0 ? -1 : 1;
    const int dmin = dl(i - d);
    int lmax = 128;  // grows by 4x (bvh.cu:208-211)
    while (dl(i + lmax * d) > dmin) lmax *= 4;
    int l = 0;
    for (int t = lmax / 2; t > 0; t /= 2)
        if (dl(i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = dl(j);
    int s = 0;
    for (int k = 2;; k *= 2) {  // t = ceil(l / k) down to 1 (bvh.cu:226-238)
        const int t = (l + k - 1) / k;
        if (dl(i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int split = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int left = (lo == split) ? split : -(split + 1);
    const int right = (hi == split + 1) ? (split + 1) : -(split + 2);
    parents[left >= 0 ? left : root_slot - (left + 1)] = i;
    parents[right >= 0 ? right : root_slot - (right + 1)] = i;
    arc_bvh_node nd;
    nd.left = left;
    nd.right = right;
    nd.bbMin = arc_float3{0.f, 0.f, 0.f};
    nd.bbMax = arc_float3{0.f, 0.f, 0.f};
    nd.visited = 0;
    nodes[i] = nd;
    ranges[2 * i] = lo;
    ranges[2 * i + 1] = hi;
    if (i == 0) parents[root_slot] = -1;
}

// ----------------------------------------------------------- refit ----
__device__ inline float len3(f3 v) {
    // sqrtf(dot(v, v)), dot evaluated left to right without fusion
    return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
}

// cluster() (bvh.cu:726-744)
__device__ inline void cluster(int target, int left, int right, int n, const NodeData& nd) {
    const int tid = n + target;
    const int lid = left < 0 ? n - (left + 1) : left;
    const int rid = right < 0 ? n - (right + 1) : right;
    const f3 rl = nd.rad[lid], rr = nd.rad[rid];
    const float i1 = len3(rl), i2 = len3(rr);
    int repr = lid;
    if (nd.rnd[tid] > (i1 / (i1 + i2))) repr = rid;  // NaN ratio keeps the left child
    nd.size[tid] = nd.size[lid] + nd.size[rid];
    nd.mat[tid] = nd.mat[repr];
    nd.rad[tid] = f3{rl.x + rr.x, rl.y + rr.y, rl.z + rr.z};
    nd.dir[tid] = nd.dir[repr];
    nd.pos[tid] = nd.pos[repr];
    nd.nrm[tid] = nd.nrm[repr];
}

// The cut's merge sums of an internal node from its children's, a (left) +
// b (right) in that order: (I, count), (sum w x, sum w), (sum w n, -).
__device__ inline void add_cut_sums(const float4 (&a)[kMrec], const float4 (&b)[kMrec], float4 (&m)[kMrec]) {
    const int count = __float_as_int(a[0].w) + __float_as_int(b[0].w);
    m[0] = make_float4(a[0].x + b[0].x, a[0].y + b[0].y, a[0].z + b[0].z, __int_as_float(count));
    m[1] = make_float4(a[1].x + b[1].x, a[1].y + b[1].y, a[1].z + b[1].z, a[1].w + b[1].w);
    m[2] = make_float4(a[2].x + b[2].x, a[2].y + b[2].y, a[2].z + b[2].z, 0.f);
}

// An internal node's cut record from its merge sums and its box lo / hi.
__device__ inline void write_cut_node(int target, int n, f3 lo, f3 hi, const float4 (&m)[kMrec], const CutRecs& cr) {
    const size_t tid = (size_t)n + target;
    const float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    const float s2 = 0.25f * (ex * ex + ey * ey + ez * ez);
    const f3 centre{0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    write_cut_record(cr.crec + kCrec * tid, f3{m[0].x, m[0].y, m[0].z}, m[1], f3{m[2].x, m[2].y, m[2].z},
                     __float_as_int(m[0].w), s2, centre, f3{0.f, 0.f, 0.f});
}

__device__ inline void store_merge(int target, int n, const float4 (&m)[kMrec], const CutRecs& cr) {
    float4* mg = cr.merge + kMrec * ((size_t)n + target);
#pragma unroll
    for (int q = 0; q < kMrec; ++q) mg[q] = m[q];
}

// The cut's sums of an internal node from its children's merge records in
// HBM, and its node record; lo / hi: the node's box.
__device__ inline void merge_cut(int target, int left, int right, int n, f3 lo, f3 hi, const CutRecs& cr) {
    const size_t lid = left < 0 ? (size_t)n - (left + 1) : (size_t)left;
    const size_t rid = right < 0 ? (size_t)n - (right + 1) : (size_t)right;
    float4 a[kMrec], b[kMrec], m[kMrec];
#pragma unroll
    for (int q = 0; q < kMrec; ++q) {
        a[q] = cr.merge[kMrec * lid + q];
        b[q] = cr.merge[kMrec * rid + q];
    }
    add_cut_sums(a, b, m);
    store_merge(target, n, m, cr);
    write_cut_node(target, n, lo, hi, m, cr);
}

// kernel_innerNodes (bvh.cu:280-316).  One thread per leaf walks up; the
// second thread to arrive at a node merges both children.  The reference
// reads the sibling's box without ordering (quirk Q3); here the arrival
// counters inside one kernel's workgroup are acq_rel RMWs at workgroup
// scope, so the first arrival's writes are visible to the second arrival
// before it reads them, and the arrivals that cross workgroups (blocks,
// spans) are relaxed: the kernel boundary publishes their writes before the
// next kernel reads them.
//
// Three kernels, no cross-XCD fence:
//   k_refit_local: a block owns kRefitLeaves consecutive leaves.  Karras node
//     k has k at one end of its leaf range, so a node whose range lies inside
//     the block's leaves has its index there too and both of its subtrees are
//     merged by this block: its counter lives in LDS (workgroup scope).  At a
//     node whose range crosses a block boundary the walk stops after a
//     relaxed increment of the node's `visited`; the thread that makes it 2
//     appends the node to the ready list.
//   k_refit_mid: one workgroup per kMidLeaves-leaf span, after the kernel
//     boundary, climbs from the ready nodes inside its span through the
//     crossing nodes inside it, with workgroup-scope counters on `visited`;
//     at a node that spans two of them it stops as k_refit_local does at a
//     block boundary, and the second arrival appends the node to a second
//     ready list.  One workgroup merging all ~10000 crossing nodes of 1M
//     points is bound by its CU's rate of scattered memory requests (about
//     20 per merge); the spans share them over tens of CUs.
//   k_refit_top: one workgroup, after the kernel boundary has published every
//     span's result, climbs from the second ready list through the remaining
//     nodes with workgroup-scope counters on `visited`.  Its width
//     is the tree's refit_top_threads: 1024 threads climb the most chains at
//     once, but a 16-wave workgroup needs 16 free wave slots on one CU, and
//     beside a kernel that fills the CUs (the gather's antiradiance on its
//     second stream) it waited for them up to 0.8 ms; the gather's trees use
//     four waves.
// Every internal node ends with visited == 2, as the reference's counters do.
constexpr int kRefitLeaves = 1024;
constexpr int kMidLeaves = 32 * kRefitLeaves, kMidThreads = 256;

// What a refit merges besides the boxes.
enum RefitMode { kBoxes = 0, kReference = 1, kCut = 2 };

template <int kMode>
__device__ inline void merge_node(int parent, arc_bvh_node* nodes, const f3* spos, int n, const NodeData& nd,
                                  const CutRecs& cr) {
    const int l = nodes[parent].left, r = nodes[parent].right;
    f3 lmin, lmax, rmin, rmax;
    if (l < 0) {
        const arc_bvh_node& c = nodes[-(l + 1)];
        lmin = f3{c.bbMin.x, c.bbMin.y, c.bbMin.z};
        lmax = f3{c.bbMax.x, c.bbMax.y, c.bbMax.z};
    } else {
        lmin = lmax = spos[l];
    }
    if (r < 0) {
        const arc_bvh_node& c = nodes[-(r + 1)];
        rmin = f3{c.bbMin.x, c.bbMin.y, c.bbMin.z};
        rmax = f3{c.bbMax.x, c.bbMax.y, c.bbMax.z};
    } else {
        rmin = rmax = spos[r];
    }
    const f3 lo{fminf(lmin.x, rmin.x), fminf(lmin.y, rmin.y), fminf(lmin.z, rmin.z)};
    const f3 hi{fmaxf(lmax.x, rmax.x), fmaxf(lmax.y, rmax.y), fmaxf(lmax.z, rmax.z)};
    nodes[parent].bbMin = arc_float3{lo.x, lo.y, lo.z};
    nodes[parent].bbMax = arc_float3{hi.x, hi.y, hi.z};
    if (kMode == kReference) cluster(parent, l, r, n, nd);
    if (kMode == kCut) merge_cut(parent, l, r, n, lo, hi, cr);
}

// A child of a block-local node: a leaf (its point, and for the light cuts'
// rebuild its merge record, from HBM) or a local node merged earlier in this
// block (from LDS).
template <int kMode>
__device__ inline void local_child(int c, int lo, const f3* spos, const CutRecs& cr, const f3* sbmin,
                                   const f3* sbmax, const float4* smerge, f3& bmin, f3& bmax, float4 (&m)[kMrec]) {
    if (c >= 0) {
        bmin = bmax = spos[c];
        if (kMode == kCut) {
#pragma unroll
            for (int q = 0; q < kMrec; ++q) m[q] = cr.merge[kMrec * (size_t)c + q];
        }
    } else {
        const int s = -(c + 1) - lo;
        bmin = sbmin[s];
        bmax = sbmax[s];
        if (kMode == kCut) {
#pragma unroll
            for (int q = 0; q < kMrec; ++q) m[q] = smerge[kMrec * s + q];
        }
    }
}

template <int kMode>
__global__ __launch_bounds__(kRefitLeaves) void k_refit_local(arc_bvh_node* __restrict__ nodes,
                                                              const int* __restrict__ parents,
                                                              const int* __restrict__ ranges,
                                                              const f3* __restrict__ spos, int n, NodeData nd,
                                                              CutRecs cr, int* __restrict__ ready,
                                                              int* __restrict__ nready) {
    // The block's local nodes, by index - lo: arrival counters, boxes and, for
    // the light cuts' rebuild, merge sums.  A local node's children are
    // leaves or local nodes, so its merge reads HBM only for leaves.
    __shared__ int arrivals[kRefitLeaves];
    __shared__ f3 sbmin[kRefitLeaves], sbmax[kRefitLeaves];
    __shared__ float4 smerge[kMode == kCut ? kMrec * kRefitLeaves : 1];
    arrivals[threadIdx.x] = 0;
    __syncthreads();
    const int lo = blockIdx.x * kRefitLeaves;
    const int hi = min(n, lo + kRefitLeaves) - 1;
    const int i = lo + threadIdx.x;
    if (i >= n) return;  // no barrier follows
    int parent = parents[i];
    bool local = parent >= 0 && ranges[2 * parent] >= lo && ranges[2 * parent + 1] <= hi;
    while (parent >= 0) {
        if (!local) {
            const int old = __hip_atomic_fetch_add(&nodes[parent].visited, 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            if (old == 1) ready[atomicAdd(nready, 1)] = parent;
            return;
        }
        if (__hip_atomic_fetch_add(&arrivals[parent - lo], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
            return;
        const int l = nodes[parent].left, r = nodes[parent].right;
        f3 lmin, lmax, rmin, rmax;
        float4 a[kMrec], b[kMrec];
        local_child<kMode>(l, lo, spos, cr, sbmin, sbmax, smerge, lmin, lmax, a);
        local_child<kMode>(r, lo, spos, cr, sbmin, sbmax, smerge, rmin, rmax, b);
        const f3 blo{fminf(lmin.x, rmin.x), fminf(lmin.y, rmin.y), fminf(lmin.z, rmin.z)};
        const f3 bhi{fmaxf(lmax.x, rmax.x), fmaxf(lmax.y, rmax.y), fmaxf(lmax.z, rmax.z)};
        const int s = parent - lo;
        sbmin[s] = blo;
        sbmax[s] = bhi;
        nodes[parent].bbMin = arc_float3{blo.x, blo.y, blo.z};
        nodes[parent].bbMax = arc_float3{bhi.x, bhi.y, bhi.z};
        if (kMode == kReference) cluster(parent, l, r, n, nd);
        const int next = parents[n + parent];
        const bool next_local = next >= 0 && ranges[2 * next] >= lo && ranges[2 * next + 1] <= hi;
        if (kMode == kCut) {
            float4 m[kMrec];
            add_cut_sums(a, b, m);
#pragma unroll
            for (int q = 0; q < kMrec; ++q) smerge[kMrec * s + q] = m[q];
            // HBM holds the merge sums the top refit reads: those of a node
            // whose parent spans blocks (and the root's)
            if (!next_local) store_merge(parent, n, m, cr);
            write_cut_node(parent, n, blo, bhi, m, cr);
        }
        nodes[parent].visited = 2;
        parent = next;
        local = next_local;
    }
}

template <int kMode>
__global__ __launch_bounds__(kMidThreads) void k_refit_mid(arc_bvh_node* __restrict__ nodes,
                                                           const int* __restrict__ parents,
                                                           const int* __restrict__ ranges,
                                                           const f3* __restrict__ spos, int n, NodeData nd,
                                                           CutRecs cr, const int* __restrict__ ready,
                                                           const int* __restrict__ nready, int* __restrict__ ready2,
                                                           int* __restrict__ nready2) {
    const int count = *nready;
    const int lo = blockIdx.x * kMidLeaves;
    const int hi = min(n, lo + kMidLeaves) - 1;
    for (int j = threadIdx.x; j < count; j += kMidThreads) {
        int node = ready[j];
        const int first = ranges[2 * node], last = ranges[2 * node + 1];
        if (first < lo || first > hi) continue;  // the span holding its first leaf takes it
        if (last > hi) {                         // spans two spans: the top refit's
            ready2[atomicAdd(nready2, 1)] = node;
            continue;
        }
        while (true) {
            merge_node<kMode>(node, nodes, spos, n, nd, cr);
            const int parent = parents[n + node];
            if (parent < 0) break;
            if (ranges[2 * parent] >= lo && ranges[2 * parent + 1] <= hi) {
                // 0: the sibling subtree is still pending and will climb on; 1: both done
                if (__hip_atomic_fetch_add(&nodes[parent].visited, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 1)
                    break;
                node = parent;
            } else {
                const int old = __hip_atomic_fetch_add(&nodes[parent].visited, 1, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                if (old == 1) ready2[atomicAdd(nready2, 1)] = parent;
                break;
            }
        }
    }
}

template <int kMode>
__global__ __launch_bounds__(kRefitLeaves) void k_refit_top(arc_bvh_node* __restrict__ nodes,
                                                            const int* __restrict__ parents,
                                                            const f3* __restrict__ spos, int n, NodeData nd,
                                                            CutRecs cr, const int* __restrict__ ready,
                                                            const int* __restrict__ nready) {
    const int count = *nready;
    for (int j = threadIdx.x; j < count; j += blockDim.x) {
        int node = ready[j];
        while (true) {
            merge_node<kMode>(node, nodes, spos, n, nd, cr);
            const int parent = parents[n + node];
            if (parent < 0) break;
            // 0: the sibling subtree is still pending and will climb on; 1: both done
            if (__hip_atomic_fetch_add(&nodes[parent].visited, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 1)
                break;
            node = parent;
        }
    }
}

inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }
// The boxes-only refit (SimpleBvh) as range queries.  A node's box is the
// min / max over the sorted leaves of its Karras range [first, last], and
// min / max are exact in any order, so the boxes equal those of the
// bottom-up merge bit for bit, without a climb or an arrival counter:
//   k_refit_points_local: one workgroup per kRefitLeaves leaves builds a
//     segment tree of their points in LDS, writes the box of each kChunk
//     leaves and of the whole block, and answers the nodes whose range lies
//     in the block (node k is one end of its range, so it lives there too);
//     the others go to a list.
//   k_refit_points_crossing: one wave per listed node: the leaves and chunks
//     at both ends of its range and the whole blocks between, read by the
//     lanes side by side and reduced across the wave.
constexpr int kChunk = 32;                             // leaves per chunk box
constexpr int kChunksPerBlock = kRefitLeaves / kChunk;  // 32

__device__ inline void grow(f3& lo, f3& hi, const f3& a, const f3& b) {
    lo = f3{fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z)};
    hi = f3{fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z)};
}

// Where the range boxes come from and go: the sorted points into the nodes'
// boxes (the SimpleBvh refit), or the sorted leaves' unit normals (the light
// cuts' leaf records) into a per-node normal box.
struct PointsIn {
    const f3* p;
    __device__ f3 operator()(int i) const { return p[i]; }
};
struct NormalsIn {
    const float4* crec;
    __device__ f3 operator()(int i) const {
        const float4 r = crec[kCrec * (size_t)i + 1];
        return f3{r.x, r.y, r.z};
    }
};
struct NodesOut {
    arc_bvh_node* nodes;
    __device__ void operator()(int k, const f3& lo, const f3& hi) const {
        nodes[k].bbMin = arc_float3{lo.x, lo.y, lo.z};
        nodes[k].bbMax = arc_float3{hi.x, hi.y, hi.z};
        nodes[k].visited = 2;
    }
};
struct BoxOut {
    float4* box;  // 2 per node: (lo, 0), (hi, 0)
    __device__ void operator()(int k, const f3& lo, const f3& hi) const {
        box[2 * (size_t)k] = make_float4(lo.x, lo.y, lo.z, 0.f);
        box[2 * (size_t)k + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
    }
};

template <class In, class Out>
__global__ __launch_bounds__(kRefitLeaves) void k_refit_points_local(Out out, const int* __restrict__ ranges,
                                                                     In spos, int n, f3* __restrict__ chunk_box,
                                                                     f3* __restrict__ block_box,
                                                                     int* __restrict__ crossing,
                                                                     int* __restrict__ ncross) {
    // segment tree: node 1 the block, nodes kRefitLeaves + t the leaves
    __shared__ f3 tlo[2 * kRefitLeaves], thi[2 * kRefitLeaves];
    __shared__ int xlist[kRefitLeaves];  // the block's nodes that cross blocks
    __shared__ int xcount, xbase;
    const int t = threadIdx.x;
    const int lo = blockIdx.x * kRefitLeaves;
    const int hi = min(n, lo + kRefitLeaves) - 1;
    const int i = lo + t;
    constexpr float kInf = __builtin_inff();
    // internal nodes are 0 .. n - 2: this thread's node's range, read while the tree is built
    const bool inner = i < n - 1;
    const int first = inner ? ranges[2 * i] : 0, last = inner ? ranges[2 * i + 1] : 0;
    if (t == 0) xcount = 0;
    if (i < n) {
        const f3 p = spos(i);
        tlo[kRefitLeaves + t] = p;
        thi[kRefitLeaves + t] = p;
    } else {  // beyond the last leaf: neutral
        tlo[kRefitLeaves + t] = f3{kInf, kInf, kInf};
        thi[kRefitLeaves + t] = f3{-kInf, -kInf, -kInf};
    }
    for (int w = kRefitLeaves / 2; w >= 1; w >>= 1) {
        __syncthreads();
        if (t < w) {
            const int k = w + t;
            f3 a = tlo[2 * k], b = thi[2 * k];
            grow(a, b, tlo[2 * k + 1], thi[2 * k + 1]);
            tlo[k] = a;
            thi[k] = b;
        }
    }
    __syncthreads();
    if (t < kChunksPerBlock && lo + t * kChunk < n) {
        const int c = blockIdx.x * kChunksPerBlock + t;
        chunk_box[2 * c] = tlo[kChunksPerBlock + t];
        chunk_box[2 * c + 1] = thi[kChunksPerBlock + t];
    }
    if (t == 0) {
        block_box[2 * blockIdx.x] = tlo[1];
        block_box[2 * blockIdx.x + 1] = thi[1];
    }
    // the crossing nodes: listed in LDS, then appended with one atomic per block
    const bool cross = inner && (first < lo || last > hi);
    if (cross) xlist[atomicAdd(&xcount, 1)] = i;
    __syncthreads();
    if (t == 0) xbase = xcount > 0 ? atomicAdd(ncross, xcount) : 0;
    __syncthreads();
    if (t < xcount) crossing[xbase + t] = xlist[t];
    if (!inner || cross) return;  // no barrier follows
    f3 blo{kInf, kInf, kInf}, bhi{-kInf, -kInf, -kInf};
    int l = first - lo + kRefitLeaves, r = last - lo + kRefitLeaves + 1;  // [l, r)
    while (l < r) {
        if (l & 1) {
            grow(blo, bhi, tlo[l], thi[l]);
            ++l;
        }
        if (r & 1) {
            --r;
            grow(blo, bhi, tlo[r], thi[r]);
        }
        l >>= 1;
        r >>= 1;
    }
    out(i, blo, bhi);
}

constexpr int kCrossThreads = 256;

template <class In, class Out>
__global__ __launch_bounds__(kCrossThreads) void k_refit_points_crossing(Out out, const int* __restrict__ ranges,
                                                                         In spos,
                                                                         const f3* __restrict__ chunk_box,
                                                                         const f3* __restrict__ block_box,
                                                                         const int* __restrict__ crossing,
                                                                         const int* __restrict__ ncross) {
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * (kCrossThreads / 64);
    const int count = *ncross;
    constexpr float kInf = __builtin_inff();
    for (int w = blockIdx.x * (kCrossThreads / 64) + (threadIdx.x >> 6); w < count; w += nwaves) {
        const int k = crossing[w];
        const int first = ranges[2 * k], last = ranges[2 * k + 1];
        // [first, last] spans blocks bf < bl, each full but the last: the
        // leaves to the end of first's chunk, the rest of block bf's chunks,
        // the blocks between, block bl's chunks before last's chunk, and the
        // leaves of last's chunk up to last
        const int bf = first / kRefitLeaves, bl = last / kRefitLeaves;
        const int cf = first / kChunk, cl = last / kChunk;
        const int nA = (cf + 1) * kChunk - first;
        const int c1 = cf + 1, nB = (bf + 1) * kChunksPerBlock - c1;
        const int b1 = bf + 1, nC = bl - b1;
        const int c3 = bl * kChunksPerBlock, nD = cl - c3;
        const int a4 = cl * kChunk, nE = last - a4 + 1;
        const int total = nA + nB + nC + nD + nE;
        f3 blo{kInf, kInf, kInf}, bhi{-kInf, -kInf, -kInf};
        for (int e = lane; e < total; e += 64) {
            f3 a, b;
            if (e < nA) {
                a = b = spos(first + e);
            } else if (e < nA + nB) {
                const int c = c1 + (e - nA);
                a = chunk_box[2 * c];
                b = chunk_box[2 * c + 1];
            } else if (e < nA + nB + nC) {
                const int q = b1 + (e - nA - nB);
                a = block_box[2 * q];
                b = block_box[2 * q + 1];
            } else if (e < nA + nB + nC + nD) {
                const int c = c3 + (e - nA - nB - nC);
                a = chunk_box[2 * c];
                b = chunk_box[2 * c + 1];
            } else {
                a = b = spos(a4 + (e - nA - nB - nC - nD));
            }
            grow(blo, bhi, a, b);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const f3 olo{__shfl_xor(blo.x, off), __shfl_xor(blo.y, off), __shfl_xor(blo.z, off)};
            const f3 ohi{__shfl_xor(bhi.x, off), __shfl_xor(bhi.y, off), __shfl_xor(bhi.z, off)};
            grow(blo, bhi, olo, ohi);
        }
        if (lane == 0) out(k, blo, bhi);
    }
}

// Chunk and block boxes of the points refit: 2 f3 per kChunk leaves and per
// kRefitLeaves leaves.
inline size_t point_boxes_bytes(int n) {
    const size_t chunks = ((size_t)n + kChunk - 1) / kChunk, blocks = ((size_t)n + kRefitLeaves - 1) / kRefitLeaves;
    return sizeof(f3) * 2 * (chunks + blocks);
}

template <class In, class Out>
void launch_range_boxes(arc_bvh* b, int n, In in, Out out, int* crossing, int* ncross) {
    const int rblocks = (n + kRefitLeaves - 1) / kRefitLeaves;
    f3* chunk_box = b->boxes.as<f3>();
    f3* block_box = chunk_box + 2 * (((size_t)n + kChunk - 1) / kChunk);
    hipLaunchKernelGGL((k_refit_points_local<In, Out>), dim3(rblocks), dim3(kRefitLeaves), 0, b->ctx->stream, out,
                       b->ranges.as<int>(), in, n, chunk_box, block_box, crossing, ncross);
    // about ten nodes per block cross blocks; a wave takes one at a time
    const int cblocks = std::max(1, std::min(1024, 2 * rblocks));
    hipLaunchKernelGGL((k_refit_points_crossing<In, Out>), dim3(cblocks), dim3(kCrossThreads), 0, b->ctx->stream,
                       out, b->ranges.as<int>(), in, chunk_box, block_box, crossing, ncross);
}

void launch_refit_points(arc_bvh* b, int n, int* crossing, int* ncross) {
    launch_range_boxes(b, n, PointsIn{b->positions.as<f3>()}, NodesOut{b->nodes.as<arc_bvh_node>()}, crossing,
                       ncross);
}


int ensure_tree(arc_bvh* b, int n) {
    arc_ctx* ctx = b->ctx;
    ARC_TRY(arc::ensure(ctx, b->codes, sizeof(uint32_t) * (size_t)n));
    ARC_TRY(arc::ensure(ctx, b->perm, sizeof(int) * (size_t)n));
    ARC_TRY(arc::ensure(ctx, b->nodes, sizeof(arc_bvh_node) * (size_t)(n - 1)));
    ARC_TRY(arc::ensure(ctx, b->parents, sizeof(int) * (size_t)(2 * n - 1)));
    ARC_TRY(arc::ensure(ctx, b->ranges, sizeof(int) * (size_t)(2 * (n - 1))));
    ARC_TRY(arc::ensure(ctx, b->positions, sizeof(f3) * (size_t)n));
    ARC_TRY(arc::ensure(ctx, b->normals, sizeof(f3) * (size_t)n));
    ARC_TRY(arc::ensure(ctx, b->bounds, sizeof(float) * 16));
    ARC_TRY(arc::ensure(ctx, b->ready, sizeof(int) * 2 * (size_t)n));  // two ready lists
    ARC_TRY(arc::ensure(ctx, b->partials, sizeof(float) * 12 * kReduceBlocks));
    ARC_TRY(arc::ensure(ctx, b->boxes, point_boxes_bytes(n)));
    return ARC_OK;
}

// Bounds + encode + sort (Bvh::create up to sort(), bvh.cu:346-374).
int encode_and_sort(arc_bvh* b, const f3* pos, const f3* nrm, int n) {
    arc_ctx* ctx = b->ctx;
    const int nparts = std::min(kReduceBlocks, std::max(1, blocks_for(n)));
    float* bounds = b->bounds.as<float>();
    hipLaunchKernelGGL(k_bounds_partial, dim3(nparts), dim3(kBlock), 0, ctx->stream, pos, nrm, n,
                       b->partials.as<float>());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(768), 0, ctx->stream, b->partials.as<float>(),
                       nparts, bounds);
    hipLaunchKernelGGL(k_encode, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, pos, nrm, n,
                       bounds, b->consider_normals && nrm ? 1 : 0, b->codes.as<uint32_t>(),
                       b->perm.as<int>());
    ARC_LAUNCHED(ctx);
    const int key_bits = (b->consider_normals && nrm) ? 32 : 30;
    return arc::sort_pairs(ctx, b->sort, b->codes.as<uint32_t>(), b->perm.as<int>(), n, key_bits);
}

template <int kMode>
void launch_refit(arc_bvh* b, int n, const NodeData& nd, const CutRecs& cr, int* ready, int* nready) {
    const int rblocks = (n + kRefitLeaves - 1) / kRefitLeaves;
    auto* nodes = b->nodes.as<arc_bvh_node>();
    hipLaunchKernelGGL(k_refit_local<kMode>, dim3(rblocks), dim3(kRefitLeaves), 0, b->ctx->stream, nodes,
                       b->parents.as<int>(), b->ranges.as<int>(), b->positions.as<f3>(), n, nd, cr, ready, nready);
    hipLaunchKernelGGL(k_refit_mid<kMode>, dim3((n + kMidLeaves - 1) / kMidLeaves), dim3(kMidThreads), 0,
                       b->ctx->stream, nodes, b->parents.as<int>(), b->ranges.as<int>(), b->positions.as<f3>(), n, nd,
                       cr, ready, nready, ready + n, nready + 1);
    hipLaunchKernelGGL(k_refit_top<kMode>, dim3(1), dim3(b->refit_top_threads), 0, b->ctx->stream, nodes,
                       b->parents.as<int>(), b->positions.as<f3>(), n, nd, cr, ready + n, nready + 1);
}

// Karras tree and refit; nd (reference node data) or cr (cut records) or
// neither (boxes only).
int build_tree(arc_bvh* b, int n, const NodeData* nd, const CutRecs* cr) {
    arc_ctx* ctx = b->ctx;
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(kBlock), 0, ctx->stream,
                       b->codes.as<uint32_t>(), n - 1, b->nodes.as<arc_bvh_node>(), b->parents.as<int>(),
                       b->ranges.as<int>(), reinterpret_cast<int*>(b->bounds.as<float>() + 13));
    int* nready = reinterpret_cast<int*>(b->bounds.as<float>() + 13);  // two counters, zeroed by k_karras
    int* ready = b->ready.as<int>();
    const NodeData d = nd ? *nd : NodeData{};
    const CutRecs c = cr ? *cr : CutRecs{};
    if (nd) launch_refit<kReference>(b, n, d, c, ready, nready);
    else if (cr) launch_refit<kCut>(b, n, d, c, ready, nready);
    else launch_refit_points(b, n, ready, nready);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

}  // namespace

namespace arc {

int random_numbers(arc_ctx* ctx, uint64_t seed, int count, float* out) {
    if (count <= 0) return ARC_OK;
    hipLaunchKernelGGL(k_random, dim3(blocks_for(count)), dim3(kBlock), 0, ctx->stream, seed, count, out);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int node_normal_boxes(arc_bvh* b, const float4* crec, float4* nbox) {
    const int n = b->n;
    if (n < 2) return ARC_OK;
    // the list and its counter of the range queries: the refit is done with them
    int* ncross = reinterpret_cast<int*>(b->bounds.as<float>() + 15);  // zeroed by k_karras
    launch_range_boxes(b, n, NormalsIn{crec}, BoxOut{nbox}, b->ready.as<int>(), ncross);
    ARC_LAUNCHED(b->ctx);
    return ARC_OK;
}

int bvh_build_points(arc_bvh* b, const f3* pos, const f3* nrm, int n) {
    if (n <= 1) {
        b->n = 0;
        return fail(b->ctx, ARC_E_TOO_FEW_POINTS, "not enough points for bvh construction");
    }
    if (b->consider_normals && !nrm) return fail(b->ctx, ARC_E_INVALID, "normals required");
    ARC_TRY(ensure_tree(b, n));
    ARC_TRY(encode_and_sort(b, pos, nrm, n));
    hipLaunchKernelGGL(k_permute_points, dim3(blocks_for(n)), dim3(kBlock), 0, b->ctx->stream,
                       b->perm.as<int>(), pos, nrm, n, b->positions.as<f3>(), b->normals.as<f3>());
    ARC_LAUNCHED(b->ctx);
    ARC_TRY(build_tree(b, n, nullptr, nullptr));
    b->n = n;
    b->has_node_data = false;
    b->has_cut_records = false;
    return ARC_OK;
}

namespace {

// Rejects bounce <= 0 (quirk Q2: such an AVPL would shift the leaves
// against the node data).
int check_bounces(arc_bvh* b, const AvplsDev& a) {
    arc_ctx* ctx = b->ctx;
    int* flag = reinterpret_cast<int*>(b->bounds.as<float>() + 12);
    ARC_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_check_bounce, dim3(blocks_for(a.n)), dim3(kBlock), 0, ctx->stream, a.bounce, a.n, flag);
    ARC_LAUNCHED(ctx);
    int host_flag = 0;
    ARC_HIP(ctx, hipMemcpyAsync(&host_flag, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (host_flag) {
        b->n = 0;
        return fail(ctx, ARC_E_INVALID, "AVPL with bounce <= 0: leaves would not match node data");
    }
    return ARC_OK;
}

}  // namespace

int bvh_build_avpls(arc_bvh* b, const AvplsDev& a, const float* rand, uint64_t seed, bool validate) {
    arc_ctx* ctx = b->ctx;
    const int n = a.n;
    if (n <= 1) {
        b->n = 0;
        return fail(ctx, ARC_E_TOO_FEW_POINTS, "not enough points for bvh construction");
    }
    ARC_TRY(ensure_tree(b, n));
    const size_t slots = (size_t)(2 * n - 1);
    ARC_TRY(ensure(ctx, b->nd_size, sizeof(int) * slots));
    ARC_TRY(ensure(ctx, b->nd_mat, sizeof(int) * slots));
    ARC_TRY(ensure(ctx, b->nd_rand, sizeof(float) * slots));
    ARC_TRY(ensure(ctx, b->nd_pos, sizeof(f3) * slots));
    ARC_TRY(ensure(ctx, b->nd_nrm, sizeof(f3) * slots));
    ARC_TRY(ensure(ctx, b->nd_rad, sizeof(f3) * slots));
    ARC_TRY(ensure(ctx, b->nd_dir, sizeof(f3) * slots));
    if (validate) ARC_TRY(check_bounces(b, a));
    // AvplBvh passes considerNormals = false from CudaGather (cudaGather.cu:858);
    // a Bvh created with normals considered keeps 6D codes here as well.
    ARC_TRY(encode_and_sort(b, a.pos, a.nrm, n));
    if (rand) {
        ARC_HIP(ctx, hipMemcpyAsync(b->nd_rand.p, rand, sizeof(float) * slots, hipMemcpyDeviceToDevice,
                                    ctx->stream));
    } else {
        ARC_TRY(random_numbers(ctx, seed, (int)slots, b->nd_rand.as<float>()));
    }
    NodeData nd{b->nd_size.as<int>(), b->nd_mat.as<int>(), b->nd_rand.as<float>(), b->nd_pos.as<f3>(),
                b->nd_nrm.as<f3>(),   b->nd_rad.as<f3>(), b->nd_dir.as<f3>()};
    ARC_TRY(ensure(ctx, b->staged, sizeof(float4) * 4 * (size_t)n));
    hipLaunchKernelGGL(k_stage_avpls, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, a, b->staged.as<float4>());
    hipLaunchKernelGGL(k_permute_avpls, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, b->perm.as<int>(), n,
                       b->staged.as<float4>(), b->positions.as<f3>(), b->normals.as<f3>(), nd);
    ARC_LAUNCHED(ctx);
    ARC_TRY(build_tree(b, n, &nd, nullptr));
    b->n = n;
    b->has_node_data = true;
    b->has_cut_records = false;
    return ARC_OK;
}

int bvh_build_cut(arc_bvh* b, const AvplsDev& a, bool validate, const arc_material* mats, float4* crec) {
    arc_ctx* ctx = b->ctx;
    const int n = a.n;
    if (n <= 1) {
        b->n = 0;
        return fail(ctx, ARC_E_TOO_FEW_POINTS, "not enough points for bvh construction");
    }
    if (!mats || !crec) return fail(ctx, ARC_E_INVALID, "cut records need the material table and a buffer");
    ARC_TRY(ensure_tree(b, n));
    const size_t slots = (size_t)(2 * n - 1);
    ARC_TRY(ensure(ctx, b->nd_merge, sizeof(float4) * kMrec * slots));
    if (validate) ARC_TRY(check_bounces(b, a));
    ARC_TRY(encode_and_sort(b, a.pos, a.nrm, n));
    const CutRecs cr{b->nd_merge.as<float4>(), crec};
    ARC_TRY(ensure(ctx, b->staged, sizeof(float4) * 4 * (size_t)n));
    hipLaunchKernelGGL(k_stage_avpls, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, a, b->staged.as<float4>());
    hipLaunchKernelGGL(k_permute_cut, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, b->perm.as<int>(), n,
                       b->staged.as<float4>(), b->positions.as<f3>(), b->normals.as<f3>(), mats, cr);
    ARC_LAUNCHED(ctx);
    ARC_TRY(build_tree(b, n, nullptr, &cr));
    b->n = n;
    b->has_node_data = false;
    b->has_cut_records = true;
    return ARC_OK;
}

}  // namespace arc
