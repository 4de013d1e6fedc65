// The moment estimator of the error-bound light cut (ARC_ESTIMATOR_MOMENTS).
//
// The reference sums a kept cluster as its random representative times the
// cluster's radiance (kernel_bvh, cudaGather.cu:416-420; cluster(),
// bvh.cu:726-744).  The centroid estimator of gather_cut.hip replaces the
// representative by the node's centroid; on scenes whose AVPLs carry
// unrelated colours and normals both leave errors of several per cent,
// because a node's clamped cosines and its intensity differ per colour
// channel and the inverse-square falloff is convex over the node.  This
// estimator keeps, per internal node:
//
//   per RGB channel c, over the node's AVPLs j (I_jc = L_jc * diffuse / pi):
//     I_c = sum I_jc,  M_c = sum I_jc n_j,  T_c = sum I_jc n_j n_j^T
//   so that the AVPL-side sum  sum_j I_jc max(0, -n_j . d^)  is taken to
//   second order in spherical harmonics: max(0, t) ~ 3/32 + t/2 + 15/32 t^2
//   (the l <= 2 Legendre terms of the clamped cosine), i.e.
//     A_c(d^) = 3/32 I_c - 1/2 M_c . d^ + 15/32 d^T T_c d^;
//   and, over w_j = (I_jr + I_jg + I_jb) / 3, the node's position mean c and
//   covariance S, from which four points c + L v_k (L L^T = S, v_k the
//   vertices (1,1,1), (1,-1,-1), (-1,1,-1), (-1,-1,1) of a tetrahedron, whose
//   mean is 0 and mean outer product the identity) share the node's mean and
//   covariance: the pixel-side factor max(0, n_p . d) / |d|^3 averaged over
//   them is exact to second order in the node's extent.
//
// The estimate of a node is sum_c max(A_c(d^), 0) * K, with d^ the direction
// from the pixel to c and K the four-point average.  The records hold A_c / 4
// (the average's 1/4, which max(., 0) lets through) with its zz term folded
// into the constant by d^ . d^ = 1:
//   A_c / 4 = a_c + b_c . d^ + (txx - tzz) x^2 + (tyy - tzz) y^2 + 2 txy xy + ...
// with a_c = (3/32 I_c + tzz) / 4, b_c = -M_c / 8 and t = 15/128 T_c; the
// pixel then sums K over the four points without scaling.  Kept nodes of at most
// cutExactCount AVPLs are summed exactly over their leaves instead (their
// sorted leaves are contiguous, crec slots [first, first + count)).
//
// The sums come from FP64 prefix sums over the sorted leaves: a node's sum is
// the difference of two prefix rows at its leaf range (the Karras ranges),
// so no tree climb is needed and the covariance is formed in FP64.
#include <algorithm>
#include <climits>
#include <cmath>

#include "cut_common.hpp"

using namespace arc::cut;

namespace {

constexpr int kMomTerms = 40;   // FP64 sums per leaf (see leaf_terms)
constexpr int kScanBlock = 256;  // leaves per block of the prefix scan
// float4 per internal node record:
//   0 (c, max_c I_c)             1 (s, first leaf, leaf count, -)
//   2-4 the four points, 12 floats   5-7 (b_c, a_c) per channel
//   8-12 per channel (txx - tzz, tyy - tzz, 0, 2txy, 2txz, 2tyz), 18 floats
constexpr int kMom = 13;
constexpr int kMomStack = 96;    // per-wave stack of the tile traversal (default; 64 and 128 also built)

// Terms of leaf i of the sorted order (crec leaf record: r0 = position,
// r1 = unit normal, r2 = I), in four groups of ten:
//   group 0: w, w x, w y, w z, w (xx, yy, zz, xy, xz, yz)   (w = mean of I,
//            x relative to the centre of the root's box: smaller prefix
//            totals, less cancellation in a node's covariance)
//   group 1 + c (channel c): I_c, I_c n (3), I_c (nx nx, ny ny, nz nz,
//   nx ny, nx nz, ny nz).
constexpr int kGroup = 10;
constexpr int kGroups = kMomTerms / kGroup;
static_assert(kGroups * kGroup == kMomTerms, "four groups of ten terms");

__device__ inline void root_centre(const arc_bvh_node* __restrict__ nodes, double* ref) {
    const arc_bvh_node& b = nodes[0];
    ref[0] = 0.5 * ((double)b.bbMin.x + b.bbMax.x);
    ref[1] = 0.5 * ((double)b.bbMin.y + b.bbMax.y);
    ref[2] = 0.5 * ((double)b.bbMin.z + b.bbMax.z);
}

__device__ inline void leaf_terms(const float4* __restrict__ crec, int i, int group, const double* ref, double* t) {
    const float4* r = crec + kCrec * (size_t)i;
    if (group == 0) {
        const float4 r0 = r[0], r2 = r[2];
        const double x = r0.x - ref[0], y = r0.y - ref[1], z = r0.z - ref[2];
        const double w = ((double)r2.x + (double)r2.y + (double)r2.z) * (1.0 / 3.0);
        t[0] = w;
        t[1] = w * x; t[2] = w * y; t[3] = w * z;
        t[4] = w * x * x; t[5] = w * y * y; t[6] = w * z * z;
        t[7] = w * x * y; t[8] = w * x * z; t[9] = w * y * z;
    } else {
        const float4 r1 = r[1], r2 = r[2];
        const double nx = r1.x, ny = r1.y, nz = r1.z;
        const double I = group == 1 ? r2.x : group == 2 ? r2.y : r2.z;
        t[0] = I;
        t[1] = I * nx; t[2] = I * ny; t[3] = I * nz;
        t[4] = I * nx * nx; t[5] = I * ny * ny; t[6] = I * nz * nz;
        t[7] = I * nx * ny; t[8] = I * nx * nz; t[9] = I * ny * nz;
    }
}

// Inclusive prefix sums of one group of terms (blockIdx.y) within each
// block of kScanBlock leaves (pre, leaf-major: 40 doubles per leaf) and the
// block totals (tot, 40 per block).
__global__ __launch_bounds__(kScanBlock) void k_mom_scan(const float4* __restrict__ crec,
                                                         const arc_bvh_node* __restrict__ nodes, int n,
                                                         double* __restrict__ pre, double* __restrict__ tot) {
    __shared__ double wt[kScanBlock / 64][kGroup];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    const int group = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v[kGroup];
    if (i < n) {
        double ref[3];
        root_centre(nodes, ref);
        leaf_terms(crec, i, group, ref, v);
    } else {
#pragma unroll
        for (int k = 0; k < kGroup; ++k) v[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double u = __shfl_up(v[k], off);
            if (lane >= off) v[k] += u;
        }
        if (lane == 63) wt[wave][k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
        double add = 0.0;
        for (int w = 0; w < wave; ++w) add += wt[w][k];
        v[k] += add;
    }
    if (i < n) {
        double* o = pre + (size_t)kMomTerms * i + kGroup * group;
#pragma unroll
        for (int k = 0; k < kGroup; ++k) o[k] = v[k];
    }
    if (threadIdx.x == kScanBlock - 1) {
        double* o = tot + (size_t)kMomTerms * blockIdx.x + kGroup * group;
#pragma unroll
        for (int k = 0; k < kGroup; ++k) o[k] = v[k];
    }
}

// Block totals -> exclusive block offsets, one workgroup per term.
__global__ __launch_bounds__(kScanBlock) void k_mom_offsets(double* __restrict__ tot, int nblocks) {
    __shared__ double wt[kScanBlock / 64];
    __shared__ double carry_s;
    const int term = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double carry = 0.0;
    for (int base = 0; base < nblocks; base += kScanBlock) {
        const int j = base + threadIdx.x;
        const double x = j < nblocks ? tot[(size_t)kMomTerms * j + term] : 0.0;
        double v = x;
        for (int off = 1; off < 64; off <<= 1) {
            const double u = __shfl_up(v, off);
            if (lane >= off) v += u;
        }
        if (lane == 63) wt[wave] = v;
        __syncthreads();
        double add = 0.0;
        for (int w = 0; w < wave; ++w) add += wt[w];
        v += add;
        if (j < nblocks) tot[(size_t)kMomTerms * j + term] = carry + v - x;
        if (threadIdx.x == kScanBlock - 1) carry_s = v;
        __syncthreads();  // every lane has read wt and carry_s is written
        carry += carry_s;
        __syncthreads();  // carry_s read before the next chunk writes it
    }
}

// Sums of terms [k0, k0 + kGroup) over sorted leaves [first, last].
__device__ inline void range_group(const double* __restrict__ pre, const double* __restrict__ tot, int first,
                                   int last, int k0, double* S) {
    const double* hi = pre + (size_t)kMomTerms * last + k0;
    const double* th = tot + (size_t)kMomTerms * (last / kScanBlock) + k0;
#pragma unroll
    for (int k = 0; k < kGroup; ++k) S[k] = hi[k] + th[k];
    if (first > 0) {
        const double* lo = pre + (size_t)kMomTerms * (first - 1) + k0;
        const double* tl = tot + (size_t)kMomTerms * ((first - 1) / kScanBlock) + k0;
#pragma unroll
        for (int k = 0; k < kGroup; ++k) S[k] -= lo[k] + tl[k];
    }
}

// One record per internal node from the prefix rows at its leaf range, one
// group of terms at a time (the record's position part, then each channel).
__global__ __launch_bounds__(kScanBlock) void k_mom_nodes(const int* __restrict__ ranges,
                                                          const arc_bvh_node* __restrict__ nodes, int n,
                                                          const double* __restrict__ pre,
                                                          const double* __restrict__ tot, float4* __restrict__ mom) {
    const int k = blockIdx.x * kScanBlock + threadIdx.x;
    if (k >= n - 1) return;
    const int first = ranges[2 * k], last = ranges[2 * k + 1];
    float4* o = mom + (size_t)kMom * k;
    double S[kGroup];
    range_group(pre, tot, first, last, 0, S);
    const arc_bvh_node& b = nodes[k];
    const double ex = (double)b.bbMax.x - b.bbMin.x, ey = (double)b.bbMax.y - b.bbMin.y,
                 ez = (double)b.bbMax.z - b.bbMin.z;
    const double s = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
    double cx, cy, cz, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
    if (S[0] > 0.0) {
        double ref[3];
        root_centre(nodes, ref);
        const double iw = 1.0 / S[0];
        const double ux = S[1] * iw, uy = S[2] * iw, uz = S[3] * iw;  // relative to ref
        sxx = fmax(S[4] * iw - ux * ux, 0.0);
        syy = fmax(S[5] * iw - uy * uy, 0.0);
        szz = fmax(S[6] * iw - uz * uz, 0.0);
        sxy = S[7] * iw - ux * uy;
        sxz = S[8] * iw - ux * uz;
        syz = S[9] * iw - uy * uz;
        cx = ref[0] + ux; cy = ref[1] + uy; cz = ref[2] + uz;
    } else {  // no intensity: the estimate is 0 wherever the point is
        cx = 0.5 * ((double)b.bbMin.x + b.bbMax.x);
        cy = 0.5 * ((double)b.bbMin.y + b.bbMax.y);
        cz = 0.5 * ((double)b.bbMin.z + b.bbMax.z);
    }
    // Cholesky factor of the covariance, pivoted: the axis of largest
    // variance first, so that the small and noisy entries of a flat or
    // collinear node (FP64 prefix differences) never divide a large one; a
    // pivot below tol ends its column
    auto var = [&](int i) -> double { return i == 0 ? sxx : i == 1 ? syy : szz; };
    int a0 = 0;
    if (syy > var(a0)) a0 = 1;
    if (szz > var(a0)) a0 = 2;
    int a1 = a0 == 0 ? 1 : 0, a2 = 3 - a0 - a1;
    if (var(a2) > var(a1)) {
        const int t = a1;
        a1 = a2;
        a2 = t;
    }
    auto cov = [&](int i, int j) -> double {  // i != j
        const int m = i + j;  // 1: xy, 2: xz, 3: yz
        return m == 1 ? sxy : m == 2 ? sxz : syz;
    };
    // (a variance below 1e-6 of the trace is prefix-difference noise: with
    // 1e6 leaves a prefix total carries about 1e-16 x 4e4 / W of it)
    const double tol = 1e-6 * (sxx + syy + szz) + 1e-300;
    double L00 = 0, L10 = 0, L20 = 0, L11 = 0, L21 = 0, L22 = 0;
    if (var(a0) > tol) {
        L00 = sqrt(var(a0));
        L10 = cov(a1, a0) / L00;
        L20 = cov(a2, a0) / L00;
        const double p11 = var(a1) - L10 * L10;
        if (p11 > tol) {
            L11 = sqrt(p11);
            L21 = (cov(a2, a1) - L20 * L10) / L11;
        }
        L22 = sqrt(fmax(var(a2) - L20 * L20 - L21 * L21, 0.0));
    }
    const double vs[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    float pts[12];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // offsets along the pivot order (a0, a1, a2), put back on x, y, z
        const double u0 = L00 * vs[q][0];
        const double u1 = L10 * vs[q][0] + L11 * vs[q][1];
        const double u2 = L20 * vs[q][0] + L21 * vs[q][1] + L22 * vs[q][2];
        const double ox = a0 == 0 ? u0 : a1 == 0 ? u1 : u2;
        const double oy = a0 == 1 ? u0 : a1 == 1 ? u1 : u2;
        const double oz = a0 == 2 ? u0 : a1 == 2 ? u1 : u2;
        pts[3 * q] = (float)(cx + ox);
        pts[3 * q + 1] = (float)(cy + oy);
        pts[3 * q + 2] = (float)(cz + oz);
    }
    o[1] = make_float4((float)s, __int_as_float(first), __int_as_float(last - first + 1), 0.f);
    o[2] = make_float4(pts[0], pts[1], pts[2], pts[3]);
    o[3] = make_float4(pts[4], pts[5], pts[6], pts[7]);
    o[4] = make_float4(pts[8], pts[9], pts[10], pts[11]);
    // per channel: A_c(d^) / 4 = 3/128 I_c + (-1/8 M_c) . d^ + d^T (15/128 T_c) d^,
    // the off-diagonal T terms doubled so that the form is a plain dot
    // product with (dx^2, dy^2, dz^2, dx dy, dx dz, dy dz), and the zz term
    // moved into the constant (dz^2 = 1 - dx^2 - dy^2 for a unit d^)
    float* tf = reinterpret_cast<float*>(o + 8);
    double Ib = 0.0;
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        range_group(pre, tot, first, last, kGroup * (1 + c), S);
        Ib = fmax(Ib, S[0]);
        const double f = 0.25 * 0.46875;
        o[5 + c] = make_float4((float)(-0.125 * S[1]), (float)(-0.125 * S[2]), (float)(-0.125 * S[3]),
                               (float)(0.25 * 0.09375 * S[0] + f * S[6]));
        tf[6 * c + 0] = (float)(f * (S[4] - S[6]));
        tf[6 * c + 1] = (float)(f * (S[5] - S[6]));
        tf[6 * c + 2] = 0.f;
        tf[6 * c + 3] = (float)(2.0 * f * S[7]);
        tf[6 * c + 4] = (float)(2.0 * f * S[8]);
        tf[6 * c + 5] = (float)(2.0 * f * S[9]);
    }
    tf[18] = 0.f;
    tf[19] = 0.f;
    o[0] = make_float4((float)cx, (float)cy, (float)cz, (float)Ib);
}

struct CutMom {
    const arc_bvh_node* nodes;  // n - 1 internal nodes (boxes)
    const float4* crec;         // leaf records, slots [0, n) in sorted order
    const float4* mom;          // kMom float4 per internal node
    int n;
};

// Wave-uniform record reads (the nodes' boxes and moment records, one
// address for the whole wave) go through the constant address space in the
// device pass: a uniform load from it is a scalar load (one request for the
// wave, the values in scalar registers) whatever the compiler can prove
// about aliasing stores.  The reads are of scalars, so no vector type's
// constructor meets the address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define ARC_UNIFORM_AS __attribute__((address_space(4)))
#else
#define ARC_UNIFORM_AS
#endif
using ufloat = const ARC_UNIFORM_AS float;
using uint_ = const ARC_UNIFORM_AS int;
__device__ inline ufloat* uptr(const void* p) { return (ufloat*)p; }
__device__ inline float4 uload4(ufloat* p) { return make_float4(p[0], p[1], p[2], p[3]); }
// A node's box and children (BvhNode layout: left, right, bbMin, bbMax, visited).
__device__ inline arc_bvh_node uload_node(const arc_bvh_node* n) {
    ufloat* f = uptr(n);
    uint_* i = (uint_*)n;
    arc_bvh_node b;
    b.left = i[0];
    b.right = i[1];
    b.bbMin = arc_float3{f[2], f[3], f[4]};
    b.bbMax = arc_float3{f[5], f[6], f[7]};
    b.visited = 0;  // not read
    return b;
}

// Squared distances carry 1e-24 (the first addend of their FMA chain) before
// their reciprocal square root: 1e36 bounds the cube of the result, at a
// point that coincides with the pixel the numerators (a cosine times |d|, or
// |d|^2) are 0, and above 1e-17 the addend is below half an ulp.
constexpr float kDistFloor2 = 1e-24f;

__device__ inline float dist2(float dx, float dy, float dz, float floor2) {
    return fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, floor2)));
}

// max(0, n_p . d) / |d|^3 at point x: the pixel-side cosine over the squared
// distance (cudaGather.cu:45-57, pixel factor), as the cosine n_p . d / |d|
// clamped to [0, 1] (the multiply's output clamp; it is at most 1 to
// rounding) times 1 / |d|^2.
__device__ inline float pixel_g(const PixelF& P, const float4& x) {
    const float dx = x.x - P.px, dy = x.y - P.py, dz = x.z - P.pz;
    const float inv = __builtin_amdgcn_rsqf(dist2(dx, dy, dz, kDistFloor2));
    const float c = __builtin_amdgcn_fmed3f((P.nx * dx + P.ny * dy + P.nz * dz) * inv, 0.f, 1.f);
    return c * (inv * inv);
}

// The node's estimate for the pixel, per channel, added to the pixel's sums.
template <bool kUniformAs>
__device__ inline void add_moments(const PixelF& P, const float4* __restrict__ r, Acc& acc) {
    // Records are read in three groups, each just before its use: fewer of
    // them are live at once in scalar registers.
    auto ld = [&](int i) -> float4 {
        if constexpr (kUniformAs) return uload4(uptr(r + i));
        else return r[i];
    };
    const float4 m0 = ld(0), p01 = ld(2), p12 = ld(3), p23 = ld(4);
    const float K = pixel_g(P, make_float4(p01.x, p01.y, p01.z, 0.f)) +
                    pixel_g(P, make_float4(p01.w, p12.x, p12.y, 0.f)) +
                    pixel_g(P, make_float4(p12.z, p12.w, p23.x, 0.f)) +
                    pixel_g(P, make_float4(p23.y, p23.z, p23.w, 0.f));
    const float dx = m0.x - P.px, dy = m0.y - P.py, dz = m0.z - P.pz;
    const float inv = __builtin_amdgcn_rsqf(dist2(dx, dy, dz, kDistFloor2));
    const float i2 = inv * inv;
    const float xx = dx * dx, yy = dy * dy, xy = dx * dy, xz = dx * dz, yz = dy * dz;
    const float4 mr = ld(5), mg = ld(6), mb = ld(7);
    const float4 t0 = ld(8), t1 = ld(9), t2 = ld(10), t3 = ld(11), t4 = ld(12);
    // (the zz coefficients t0.z, t2.x and t3.z are 0: folded into the constants)
    const float qr = t0.x * xx + t0.y * yy + t0.w * xy + t1.x * xz + t1.y * yz;
    const float qg = t1.z * xx + t1.w * yy + t2.y * xy + t2.z * xz + t2.w * yz;
    const float qb = t3.x * xx + t3.y * yy + t3.w * xy + t4.x * xz + t4.y * yz;
    const float ar = mr.w + inv * (mr.x * dx + mr.y * dy + mr.z * dz) + i2 * qr;
    const float ag = mg.w + inv * (mg.x * dx + mg.y * dy + mg.z * dz) + i2 * qg;
    const float ab = mb.w + inv * (mb.x * dx + mb.y * dy + mb.z * dz) + i2 * qb;
    acc.s0 += fmaxf(ar, 0.f) * K;
    acc.s1 += fmaxf(ag, 0.f) * K;
    acc.s2 += fmaxf(ab, 0.f) * K;
    ++acc.run;
}

// G of one leaf (an AVPL with a unit normal n_a) for the pixel:
// max(0, n . d) max(0, -n_a . d) / |d|^4 (cudaGather.cu:45-57), from one
// reciprocal of |d|^2.  With unit normals the cosines need no clamp at 1;
// |d|^2 carries 1e-18 (below half an ulp from 1e-11 up), and coincident
// points give 0, as both cosines vanish there.
__device__ inline float leaf_g(const PixelF& P, const float4& r0, const float4& r1) {
    const float dx = r0.x - P.px, dy = r0.y - P.py, dz = r0.z - P.pz;
    const float dd = dist2(dx, dy, dz, 1e-18f);
    const float u1 = fmaxf(P.nx * dx + P.ny * dy + P.nz * dz, 0.f);
    const float u2 = fmaxf(-(r1.x * dx + r1.y * dy + r1.z * dz), 0.f);
    const float r = __builtin_amdgcn_rcpf(dd);
    return (u1 * u2) * (r * r);
}

// The Lightcuts bound of a node for the pixel (node_bound) with one
// transcendental: 1 / dmin^2 as the square of 1 / dmin.  The cosine bound
// min(1, hmax / dmin), 0 where hmax <= 0, is hmax / dmin clamped to [0, 1]
// (the multiply's output clamp).
__device__ inline float node_bound1(const PixelF& P, const arc_bvh_node& b, float ilum, float delta2) {
    const float ex = fmaxf(fmaxf(b.bbMin.x - P.px, P.px - b.bbMax.x), 0.f);
    const float ey = fmaxf(fmaxf(b.bbMin.y - P.py, P.py - b.bbMax.y), 0.f);
    const float ez = fmaxf(fmaxf(b.bbMin.z - P.pz, P.pz - b.bbMax.z), 0.f);
    const float dmin2 = fmaxf(ex * ex + ey * ey + ez * ez, delta2);
    const float hmax = P.nx * (P.nx > 0.f ? b.bbMax.x - P.px : b.bbMin.x - P.px) +
                       P.ny * (P.ny > 0.f ? b.bbMax.y - P.py : b.bbMin.y - P.py) +
                       P.nz * (P.nz > 0.f ? b.bbMax.z - P.pz : b.bbMin.z - P.pz);
    const float ri = __builtin_amdgcn_rsqf(dmin2);
    return ilum * __builtin_amdgcn_fmed3f(hmax * ri, 0.f, 1.f) * (ri * ri);
}

// Inclusive sum over the wave's lanes.
__device__ inline int wave_inclusive_sum(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// The queued exact ranges' leaves, summed exactly.  A range is
// (first << 7 | (count - 1)) over contiguous sorted leaves.  The leaves of
// all ranges are taken 64 at a time: each lane finds its leaf (a binary
// search over the ranges' prefix counts) and loads its record with vector
// loads, the wave writes the 64 records to its LDS stage, and the pairs are
// evaluated from there, every lane reading the same record (a broadcast).
// The next 64 records are loaded while these are evaluated, so a record's
// memory latency is paid once per 64 leaves, not once per leaf.
//
// A leaf whose AVPL faces away from every pixel of the tile adds 0 to every
// sum (its cosine max(0, n_a . (p - x_a)) is 0), and about half of the
// leaves do (a surface's AVPLs seen from behind): each lane tests its leaf
// against the tile's box tb = (centre, half extent): the largest
// n_a . (p - x_a) over the box is n_a . (c - x_a) + |n_a| . e.  Only the
// leaves that can face some pixel, with a margin above FP32 rounding, are
// staged.
__device__ inline void add_ranges(const float4* __restrict__ crec, const uint32_t* X, int nx, int* pfx,
                                  float4 (*stage)[3], const float* tb, const PixelF& P, Acc& acc) {
    const int lane = threadIdx.x & 63;
    const int c0 = lane < nx ? (int)(X[lane] & 127u) + 1 : 0;
    const int c1 = lane + 64 < nx ? (int)(X[lane + 64] & 127u) + 1 : 0;
    const int s0 = wave_inclusive_sum(c0);
    const int t0 = __builtin_amdgcn_readlane(s0, 63);
    const int s1 = wave_inclusive_sum(c1) + t0;
    const int total = __builtin_amdgcn_readlane(s1, 63);
    if (lane < nx) pfx[lane] = s0 - c0;
    if (lane + 64 < nx) pfx[lane + 64] = s1 - c1;
    // the wave's LDS writes come before its reads of other lanes' slots
    __builtin_amdgcn_wave_barrier();
    auto fetch = [&](int base, float4& a0, float4& a1, float4& a2) {
        const int g = base + lane;
        if (g < total) {
            int lo = 0, hi = nx - 1;  // the last range that starts at or before g
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pfx[mid] <= g) lo = mid;
                else hi = mid - 1;
            }
            const float4* l = crec + kCrec * (size_t)((int)(X[lo] >> 7) + (g - pfx[lo]));
            a0 = l[0];
            a1 = l[1];
            a2 = l[2];
        }
    };
    float4 b0 = make_float4(0.f, 0.f, 0.f, 0.f), b1 = b0, b2 = b0;
    fetch(0, b0, b1, b2);
    for (int base = 0; base < total; base += 64) {
        bool facing = base + lane < total;
        if (facing) {
            const float dx = tb[0] - b0.x, dy = tb[1] - b0.y, dz = tb[2] - b0.z;
            const float reach = b1.x * dx + b1.y * dy + b1.z * dz + fabsf(b1.x) * tb[3] + fabsf(b1.y) * tb[4] +
                                fabsf(b1.z) * tb[5];
            const float margin = 1e-5f * (fabsf(b0.x) + fabsf(b0.y) + fabsf(b0.z) + fabsf(tb[0]) + fabsf(tb[1]) +
                                          fabsf(tb[2]) + tb[3] + tb[4] + tb[5]);
            facing = reach > -margin;
        }
        const uint64_t fm = __ballot(facing);
        if (facing) {
            const int slot = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0u));
            stage[slot][0] = b0;
            stage[slot][1] = b1;
            stage[slot][2] = b2;
        }
        __builtin_amdgcn_wave_barrier();
        if (base + 64 < total) fetch(base + 64, b0, b1, b2);
        const int cnt = __popcll(fm);
        // not unrolled: unrolled, the loop holds several leaves' records in
        // registers and the kernel loses waves per SIMD (measured slower)
#pragma unroll 1
        for (int j = 0; j < cnt; ++j) {
            const float4 r0 = stage[j][0], r1 = stage[j][1], r2 = stage[j][2];
            acc.add_run(r2, leaf_g(P, r0, r1));
        }
        acc.maybe_flush();
        __builtin_amdgcn_wave_barrier();  // every lane's reads before the next chunk's writes
    }
}

// One depth-first cut per 8x8 tile (one wave), refined when any of the
// tile's pixels asks for it (ballot), as k_cut_eb_tile; kept nodes are summed
// with the moment estimate, or exactly when they hold at most exact_count
// AVPLs.  criterion: ARC_CRITERION_BOUND (bound > threshold) or
// ARC_CRITERION_EXPANSION (s > opening * d_c, or bound * (s / d_c)^2 >
// threshold).
//
// The traversal only decides: kept nodes and exact leaf ranges are queued in
// LDS (kKeep of each per wave) and summed in passes of their own when a queue
// fills and at the end.  The traversal loop then carries no sums across its
// branches, and each pass is a straight loop of one kind of work.
constexpr int kKeep = 128;

// Kept nodes handed to the matrix-core pass (k_mom_est) instead of being
// estimated in the traversal: a frame's pool of chunks of kKeep node
// indices, taken with one atomic per chunk, each chunk linked to the
// previous chunk of its tile.  A tile whose chunk does not fit estimates it
// in place, so the pool's size changes only the speed.
struct EstPool {
    uint32_t* entries;  // cap chunks of kKeep internal-node indices
    int2* meta;         // per chunk: (entries, the tile's previous chunk or -1)
    int* taken;         // chunks taken this frame (may pass cap)
    int cap;
    int* head;          // per tile: its last chunk, -1 for none
    float* queued;      // per tile: the nodes it queued (the order of k_mom_est)
};

// The cut of one 8x8 tile (tx, ty) of a view, by one wave, with the wave's
// LDS: stack S, the two queues E and X, the exact ranges' prefix counts and
// the leaf stage.  The tree's arrays are read-only, non-aliasing arguments:
// the compiler then reads the wave-uniform node and moment records with
// scalar loads.
// The traversal's node and record reads: through the constant address space
// (kUniformAs, a copy of the node) or plainly (a reference into the array,
// so fields are read where they are used).
template <bool kUniformAs>
struct NodeRead;
template <>
struct NodeRead<false> {
    __device__ static const arc_bvh_node& node(const arc_bvh_node* __restrict__ n, int k) { return n[k]; }
    __device__ static float4 rec(const float4* __restrict__ m, int k, int i) { return m[kMom * (size_t)k + i]; }
};
template <>
struct NodeRead<true> {
    __device__ static arc_bvh_node node(const arc_bvh_node* n, int k) { return uload_node(n + k); }
    __device__ static float4 rec(const float4* m, int k, int i) { return uload4(uptr(m + kMom * (size_t)k + i)); }
};

template <bool kUniformAs, bool kPool, int kStack>
__device__ __forceinline__ void cut_tile(const arc_bvh_node* __restrict__ nodes, const float4* __restrict__ crec,
                                         const float4* __restrict__ mom, const float4* __restrict__ nbox,
                                         const float4* __restrict__ gpos,
                                         const float4* __restrict__ gnrm, const float4* __restrict__ galb,
                                         const arc_material* __restrict__ mats, int W, int H, const Threshold& th,
                                         float delta, int budget, int criterion, float open2, int exact_count,
                                         float* __restrict__ tile_cost, float4* __restrict__ out_rad,
                                         unsigned long long* __restrict__ counters, int tx, int ty, int tw, int* S,
                                         int* E, uint32_t* X, int* pfx, float4 (*stage)[3], float* tb,
                                         const EstPool* __restrict__ pool) {
    const int lane = threadIdx.x & 63;
    const int2 xy = make_int2(tx * 8 + (lane & 7), ty * 8 + (lane >> 3));
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    const float thr = live ? th.at(q) : INFINITY;
    const float delta2 = delta * delta;
    {  // the box of the tile's live pixels, as centre and half extent, for the exact leaves
        float lx = live ? P.px : INFINITY, ly = live ? P.py : INFINITY, lz = live ? P.pz : INFINITY;
        float hx = live ? P.px : -INFINITY, hy = live ? P.py : -INFINITY, hz = live ? P.pz : -INFINITY;
        for (int off = 32; off > 0; off >>= 1) {
            lx = fminf(lx, __shfl_xor(lx, off));
            ly = fminf(ly, __shfl_xor(ly, off));
            lz = fminf(lz, __shfl_xor(lz, off));
            hx = fmaxf(hx, __shfl_xor(hx, off));
            hy = fmaxf(hy, __shfl_xor(hy, off));
            hz = fmaxf(hz, __shfl_xor(hz, off));
        }
        if (lane == 0) {
            tb[0] = 0.5f * (lx + hx);
            tb[1] = 0.5f * (ly + hy);
            tb[2] = 0.5f * (lz + hz);
            tb[3] = 0.5f * (hx - lx);
            tb[4] = 0.5f * (hy - ly);
            tb[5] = 0.5f * (hz - lz);
        }
        __builtin_amdgcn_wave_barrier();  // lane 0's writes before the wave's reads
    }
    const bool expansion = criterion == ARC_CRITERION_EXPANSION;
    Acc acc;
    unsigned nvis = 0, nev = 0, nerr = 0, nbud = 0, ntest = 0, nexact = 0;  // wave-uniform
    int ne = 0, nx = 0;                                                     // queued, wave-uniform
    int head = -1;                                                          // the tile's last pool chunk
    unsigned npool = 0;                                                     // nodes handed to the pool
    // the queued kept nodes: to the pool when there is room, else estimated here
    auto flush_est = [&]() {
        if (ne == 0) return;
        // A kept node whose AVPLs all face away from every pixel of the tile
        // adds 0 (its estimate would add only the clamped cosine's
        // second-order residue): the largest n_a . (p - x_a) over the
        // node's normal box, its box and the tile's box, by interval
        // products per axis, is below 0 by more than FP32 rounding.  One
        // lane per queued node; the facing ones are compacted in place.
        int kept = 0;
        for (int base = 0; base < ne; base += 64) {
            const int j = base + lane;
            bool facing = false;
            int k = 0;
            if (j < ne) {
                k = E[j];
                const arc_bvh_node& nd = nodes[k];
                const float4 nl = nbox[2 * (size_t)k], nh = nbox[2 * (size_t)k + 1];
                float reach = 0.f, scale = 0.f;
                auto axis = [&](float c, float e, float plo, float phi, float nlo, float nhi) {
                    const float vlo = (c - e) - phi, vhi = (c + e) - plo;  // p - x over the boxes
                    reach += fmaxf(fmaxf(nlo * vlo, nlo * vhi), fmaxf(nhi * vlo, nhi * vhi));
                    scale += fabsf(c) + e + fabsf(plo) + fabsf(phi);
                };
                axis(tb[0], tb[3], nd.bbMin.x, nd.bbMax.x, nl.x, nh.x);
                axis(tb[1], tb[4], nd.bbMin.y, nd.bbMax.y, nl.y, nh.y);
                axis(tb[2], tb[5], nd.bbMin.z, nd.bbMax.z, nl.z, nh.z);
                facing = reach > -1e-5f * scale;
            }
            const uint64_t fm = __ballot(facing);
            // every lane has read its entry before any is overwritten (kept <= base)
            __builtin_amdgcn_wave_barrier();
            if (facing)
                E[kept + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(fm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)fm, 0u))] = k;
            kept += __popcll(fm);
            __builtin_amdgcn_wave_barrier();
        }
        nev -= (unsigned)(ne - kept);
        ne = kept;
        if (ne == 0) return;
        if constexpr (kPool) {
            int c = 0;
            if (lane == 0) c = atomicAdd(pool->taken, 1);
            c = uniform(c);  // lane 0's
            if (c < pool->cap) {
                uint32_t* dst = pool->entries + (size_t)c * kKeep;
                if (lane < ne) dst[lane] = (uint32_t)E[lane];
                if (lane + 64 < ne) dst[lane + 64] = (uint32_t)E[lane + 64];
                if (lane == 0) pool->meta[c] = make_int2(ne, head);
                head = c;
                npool += ne;
                ne = 0;
                return;
            }
        }
        for (int j = 0; j < ne; ++j) {
            const int k = uniform(E[j]);
            add_moments<kUniformAs>(P, mom + kMom * (size_t)k, acc);
            acc.maybe_flush();
        }
        ne = 0;
    };
    auto drain_exact = [&]() {
        if (nx > 0) add_ranges(crec, X, nx, pfx, stage, tb, P, acc);
        nx = 0;
    };
    int sp = 0;
    int node = -1;
    for (;;) {
        bool pop = true;  // take the next node from the stack (not a refined node's left child)
        if (node < 0) {
            const int k = -(node + 1);
            const auto& b = NodeRead<kUniformAs>::node(nodes, k);
            const float4 m0 = NodeRead<kUniformAs>::rec(mom, k, 0);  // (c, max I)
            const float4 m1 = NodeRead<kUniformAs>::rec(mom, k, 1);  // (s, first, count, -)
            const float s = m1.x;
            const int first = __float_as_int(m1.y), count = __float_as_int(m1.z);
            const float bound = node_bound1(P, b, m0.w, delta2);
            if (!__ballot(bound > 0.f)) {
                // behind every pixel's tangent plane: contributes exactly 0
            } else if (count <= exact_count) {
                // small node: its AVPLs exactly (refining it would end in the same sum)
                X[nx++] = ((uint32_t)first << 7) | (uint32_t)(count - 1);
                nexact += count;
            } else {
                bool refine;
                if (expansion) {
                    const float dx = m0.x - P.px, dy = m0.y - P.py, dz = m0.z - P.pz;
                    const float dc2 = dx * dx + dy * dy + dz * dz;
                    const float s2 = s * s;
                    refine = bound > 0.f && (s2 > open2 * dc2 || bound * s2 > thr * dc2);
                } else {
                    refine = bound > thr;
                }
                ++ntest;
                bool keep = true;
                if (__ballot(refine)) {
                    if ((int)(nev + nexact) >= budget) {
                        ++nbud;  // cut budget spent: keep the cluster
                    } else if (sp < kStack) {
                        S[sp++] = b.right;
                        node = uniform(b.left);
                        ++nvis;
                        keep = false;
                        pop = false;
                    } else {
                        ++nerr;  // stack full: keep the cluster
                    }
                }
                if (keep) {
                    E[ne++] = k;
                    ++nev;
                }
            }
        } else {
            X[nx++] = (uint32_t)node << 7;  // a leaf: a range of one
            ++nexact;
        }
        if (ne == kKeep) flush_est();
        if (nx == kKeep) drain_exact();
        if (pop) {
            if (sp == 0) break;
            node = uniform(S[--sp]);
        }
    }
    flush_est();
    drain_exact();
    acc.flush();
    // the tile's cost in VALU instructions (about 30 per bound test, 130 per
    // moment estimate made here, 18 per queued exact leaf: 26 per pair for
    // the half that face the tile, a few for the rest; and the pixel's own
    // loads and stores): the next frame's launch order
    if (lane == 0) {
        tile_cost[ty * tw + tx] = 300.f + 30.f * (float)ntest + 130.f * (float)(nev - npool) +
                                  18.f * (float)nexact + 2.f * (float)npool;
        if constexpr (kPool) {
            pool->head[ty * tw + tx] = head;
            pool->queued[ty * tw + tx] = (float)npool;
        }
    }
    const int alive = __popcll(__ballot(live));  // all lanes take part in the ballot
    if (counters && lane == 0) {
        atomicAdd(&counters[C_TESTED], (unsigned long long)ntest * alive);
        atomicAdd(&counters[C_VISITED], (unsigned long long)nvis * alive);
        atomicAdd(&counters[C_EVALUATED], (unsigned long long)(nev + nexact) * alive);
        atomicAdd(&counters[C_EXACT], (unsigned long long)nexact * alive);
        atomicAdd(&counters[C_ERRORS], (unsigned long long)nerr);
        atomicAdd(&counters[C_BUDGET], (unsigned long long)nbud);
    }
    if (live)
        out_rad[q] = make_float4(acc.k[0].s * P.br, acc.k[1].s * P.bg, acc.k[2].s * P.bb,
                                 counters ? (float)(nev + nexact) : 1.f);
}

// One launch per view: the view's tiles, in `order` when given (heaviest
// first), else each block its four tiles of a 16x16-pixel block.
//
// Eight waves per SIMD: the registers are held to 64 VGPRs (a few values of
// the tile's set-up spill, none inside the traversal or the sums), and a
// block's LDS (about 20 KB with a stack of 64 or 96 nodes) lets eight blocks
// share a CU's 160 KB.  The traversal waits on one dependent node load at a
// time, so the eighth wave hides more of it: 37.5 -> 35.8 ms on the 1080p x
// 1M frame.
#define ARC_CUT_WAVES __attribute__((amdgpu_waves_per_eu(8)))
template <bool kPool, int kStack>
__global__ __launch_bounds__(kCutThreads) ARC_CUT_WAVES void k_cut_mom_tile(const arc_bvh_node* __restrict__ nodes,
                                                              const float4* __restrict__ crec,
                                                              const float4* __restrict__ mom,
                                                              const float4* __restrict__ nbox,
                                                              const float4* __restrict__ gpos,
                                                              const float4* __restrict__ gnrm,
                                                              const float4* __restrict__ galb,
                                                              const arc_material* __restrict__ mats, int W, int H,
                                                              Threshold th, float delta, int budget, int criterion,
                                                              float open2, int exact_count,
                                                              const int* __restrict__ order,
                                                              float* __restrict__ tile_cost,
                                                              float4* __restrict__ out_rad,
                                                              unsigned long long* __restrict__ counters,
                                                              const EstPool* __restrict__ pool) {
    __shared__ int stk[kCutThreads / 64][kStack];
    __shared__ int qest[kCutThreads / 64][kKeep];       // internal nodes to estimate
    __shared__ uint32_t qex[kCutThreads / 64][kKeep];   // exact ranges: first << 7 | (count - 1)
    __shared__ int qpfx[kCutThreads / 64][kKeep];       // their leaves' prefix counts
    __shared__ float4 lstage[kCutThreads / 64][64][3];  // 64 leaf records
    __shared__ float tbox[kCutThreads / 64][8];         // the tile's box: centre, half extent
    const int wave = threadIdx.x >> 6;
    // the wave's 8x8 tile: the tile grid is (2 gx) x (2 gy) tiles of the
    // launch's gx x gy blocks of 16x16 pixels
    const int gx = (W + 15) / 16, gy = (H + 15) / 16, tw = 2 * gx;
    int tx, ty;
    if (order) {
        const int ti = uniform(order[blockIdx.x * 4 + wave]);
        tx = ti % tw;
        ty = ti / tw;
    } else {
        const int b = xcd_block(blockIdx.x, gx * gy);
        tx = (b % gx) * 2 + (wave & 1);
        ty = (b / gx) * 2 + (wave >> 1);
    }
    cut_tile<false, kPool, kStack>(nodes, crec, mom, nbox, gpos, gnrm, galb, mats, W, H, th, delta, budget, criterion, open2,
                           exact_count, tile_cost, out_rad, counters, tx, ty, tw, stk[wave], qest[wave], qex[wave],
                           qpfx[wave], lstage[wave], tbox[wave], pool);
}

// One launch over the tiles of many views: the launch's tiles are the views'
// tile lists one after the other (view v's from views[v].tile_base), taken
// in `order` when given, so the heaviest tiles of all views start first and
// no view's tail runs alone.
__global__ __launch_bounds__(kCutThreads) ARC_CUT_WAVES void k_cut_mom_views(const arc::CutView* __restrict__ views, int nviews,
                                                               const int* __restrict__ order, int nitems) {
    __shared__ int stk[kCutThreads / 64][kMomStack];
    __shared__ int qest[kCutThreads / 64][kKeep];
    __shared__ uint32_t qex[kCutThreads / 64][kKeep];
    __shared__ int qpfx[kCutThreads / 64][kKeep];
    __shared__ float4 lstage[kCutThreads / 64][64][3];
    __shared__ float tbox[kCutThreads / 64][8];
    const int wave = threadIdx.x >> 6;
    const int slot = uniform(blockIdx.x * 4 + wave);  // one per wave (the compiler cannot see that)
    if (slot >= nitems) return;  // the whole wave: the cut has no block barrier
    const int item = order ? uniform(order[slot]) : slot;
    int lo = 0, hi = nviews - 1;  // the last view whose tiles start at or before item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].tile_base <= item) lo = mid;
        else hi = mid - 1;
    }
    const arc::CutView& v = views[lo];
    const int t = item - v.tile_base;
    const int tw = 2 * ((v.W + 15) / 16);
    Threshold th;
    th.est = v.th_est;
    th.sum = v.th_sum;
    th.eps = v.th_eps;
    th.floor_frac = v.th_floor;
    th.inv_npix = v.th_inv_npix;
    cut_tile<true, false, kMomStack>(v.nodes, v.crec, v.mom, v.nbox, v.gpos, v.gnrm, v.galb, v.mats, v.W, v.H, th, v.delta, v.budget,
                          v.criterion, v.open2, v.exact_count, v.tile_cost, v.out_rad, v.counters, t % tw, t / tw, tw,
                          stk[wave], qest[wave], qex[wave], qpfx[wave], lstage[wave], tbox[wave], nullptr);
}

// ---- kept nodes on the matrix cores ---------------------------------------
//
// A kept node's estimate for a pixel p (add_moments) is made of dot products
// of pixel features with node coefficients.  With d = c - p:
//   |d|^2              = |c|^2 - 2 c . p + |p|^2
//   n_p . (q_j - p)    = (q_j, -1) . (n_p, n_p . p)          (the four points)
//   |q_j - p|^2        = |q_j|^2 - 2 q_j . p + |p|^2
//   b_c . d            = (b_c . c, -b_c) . (1, p)
//   d^T T'_c d         = a quadratic polynomial in p: 10 coefficients,
// where T'_c is the record's quadratic form plus a_c on its diagonal, which
// folds the constant term in (a_c = a_c |d|^2 / |d|^2): 15/128 T_c plus
// 3/128 I_c on the diagonal.  Per channel
//   A_c / 4 = b_c . d / |d| + d^T T'_c d / |d|^2,
// and the node adds max(A_c / 4, 0) * sum_j max(n_p . (q_j - p), 0) / |q_j - p|^3.
// For 16 nodes x 16 pixels each product is one 16x16x4 FP32 MFMA (the
// quadratic form three, K = 12): 21 per block.  What stays on the VALU is
// the five reciprocal square roots, the clamps and the sums.
//
// Coordinates are relative to the tile's first pixel.  The rounding of an
// expanded square |c|^2 - 2 c . p + |p|^2 is relative to |c|^2 + |p|^2,
// which stays within a small factor of |d|^2 for a kept node (its bound is
// small: the node is far against its own size and the tile's).
//
// Lane l of v_mfma_f32_16x16x4f32 holds A[l % 16][l / 16], B[l / 16][l % 16]
// and D[4 (l / 16) + r][l % 16], r = 0..3.  A row is a node (16 per group),
// a column a pixel (16 per block, four blocks per tile), so lane l of the
// result holds pixel l % 16 of the block for nodes 4 (l / 16) .. + 3: its sums
// run over a quarter of the nodes and meet across lanes once at the end.
typedef float mf4 __attribute__((ext_vector_type(4)));

__device__ inline mf4 mfma4(float a, float b, mf4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

constexpr int kEstRows = 21;  // MFMA A operands per node
constexpr int kEstFeat = 17;  // pixel features (16 in four sets of 4, and |p|^2)

// Row k of the A operands of node record r (kMom float4), relative to o.
__device__ inline void est_rows(const float4* __restrict__ r, float ox, float oy, float oz, int k,
                                float (&a)[kEstRows]) {
    const float4 m0 = r[0];
    const float cx = m0.x - ox, cy = m0.y - oy, cz = m0.z - oz;
    // component k - 1 of a vector (rows 1..3)
    auto comp = [k](float x, float y, float z) { return k == 1 ? x : (k == 2 ? y : z); };
    a[0] = k == 0 ? cx * cx + cy * cy + cz * cz : -2.f * comp(cx, cy, cz);
    const float4 p01 = r[2], p12 = r[3], p23 = r[4];
    const float q[12] = {p01.x, p01.y, p01.z, p01.w, p12.x, p12.y, p12.z, p12.w, p23.x, p23.y, p23.z, p23.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float qx = q[3 * j] - ox, qy = q[3 * j + 1] - oy, qz = q[3 * j + 2] - oz;
        a[1 + 2 * j] = k == 0 ? qx : (k == 1 ? qy : (k == 2 ? qz : -1.f));
        a[2 + 2 * j] = k == 0 ? qx * qx + qy * qy + qz * qz : -2.f * comp(qx, qy, qz);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float4 mr = r[5 + ch];
        const float4 ta = r[8 + (6 * ch) / 4], tb = r[8 + (6 * ch + 4) / 4];
        // the six floats 6 ch .. 6 ch + 5 of records 8..12: (xx, yy, zz, 2xy, 2xz, 2yz)
        const float tt[8] = {ta.x, ta.y, ta.z, ta.w, tb.x, tb.y, tb.z, tb.w};
        const int o6 = (6 * ch) % 4;
        const float al = mr.w;
        const float txx = tt[o6] + al, tyy = tt[o6 + 1] + al, tzz = tt[o6 + 2] + al;
        const float txy = tt[o6 + 3], txz = tt[o6 + 4], tyz = tt[o6 + 5];
        a[9 + ch] = k == 0 ? mr.x * cx + mr.y * cy + mr.z * cz : -comp(mr.x, mr.y, mr.z);
        const float lx = -(2.f * txx * cx + txy * cy + txz * cz);
        const float ly = -(2.f * tyy * cy + txy * cx + tyz * cz);
        const float lz = -(2.f * tzz * cz + txz * cx + tyz * cy);
        const float c0 = txx * cx * cx + tyy * cy * cy + tzz * cz * cz + txy * cx * cy + txz * cx * cz + tyz * cy * cz;
        a[12 + 3 * ch] = k == 0 ? c0 : comp(lx, ly, lz);
        a[13 + 3 * ch] = k == 0 ? txx : (k == 1 ? tyy : (k == 2 ? tzz : txy));
        a[14 + 3 * ch] = k == 0 ? txz : (k == 1 ? tyz : 0.f);
    }
}

// The pooled kept nodes of each tile, one wave per tile, added to the
// traversal's sums in out_rad.  Tiles in `order` when given.
__global__ __launch_bounds__(kCutThreads) void k_mom_est(const float4* __restrict__ mom,
                                                         const uint32_t* __restrict__ entries,
                                                         const int2* __restrict__ meta,
                                                         const int* __restrict__ head,
                                                         const float4* __restrict__ gpos,
                                                         const float4* __restrict__ gnrm,
                                                         const float4* __restrict__ galb,
                                                         const arc_material* __restrict__ mats, int W, int H,
                                                         const int* __restrict__ order,
                                                         float4* __restrict__ out_rad) {
    __shared__ float feat[kCutThreads / 64][kEstFeat][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16, tw = 2 * gx;
    int tx, ty;
    if (order) {
        const int ti = uniform(order[blockIdx.x * 4 + wave]);
        tx = ti % tw;
        ty = ti / tw;
    } else {
        const int b = xcd_block(blockIdx.x, gx * gy);
        tx = (b % gx) * 2 + (wave & 1);
        ty = (b / gx) * 2 + (wave >> 1);
    }
    int c = uniform(head[ty * tw + tx]);
    if (c < 0) return;  // the whole wave: nothing pooled (no block barrier below)
    const int2 xy = make_int2(tx * 8 + (lane & 7), ty * 8 + (lane >> 3));
    const bool live = xy.x < W && xy.y < H;
    const int q = live ? xy.y * W + xy.x : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    // the tile's origin: its first pixel (lane 0's)
    const float o_x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(P.px)));
    const float o_y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(P.py)));
    const float o_z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(P.pz)));
    {
        const float px = P.px - o_x, py = P.py - o_y, pz = P.pz - o_z;
        float (*f)[64] = feat[wave];
        f[0][lane] = 1.f;
        f[1][lane] = px;
        f[2][lane] = py;
        f[3][lane] = pz;
        f[4][lane] = P.nx;
        f[5][lane] = P.ny;
        f[6][lane] = P.nz;
        f[7][lane] = P.nx * px + P.ny * py + P.nz * pz;
        f[8][lane] = px * px;
        f[9][lane] = py * py;
        f[10][lane] = pz * pz;
        f[11][lane] = px * py;
        f[12][lane] = px * pz;
        f[13][lane] = py * pz;
        f[14][lane] = 0.f;
        f[15][lane] = 0.f;
        f[16][lane] = px * px + py * py + pz * pz;
    }
    __builtin_amdgcn_wave_barrier();  // the wave's feature writes before its reads of other lanes'
    const int kr = lane >> 4, col = lane & 15;
    const mf4 zero = {0.f, 0.f, 0.f, 0.f};
    float tot[4][3];
#pragma unroll
    for (int b = 0; b < 4; ++b) tot[b][0] = tot[b][1] = tot[b][2] = 0.f;
    while (c >= 0) {
        const int2 m = meta[c];
        const int cnt = uniform(m.x), prev = uniform(m.y);
        for (int g0 = 0; g0 < cnt; g0 += 16) {
            float a[kEstRows];
            if (g0 + col < cnt) {
                const int node = (int)entries[(size_t)c * kKeep + g0 + col];
                est_rows(mom + kMom * (size_t)node, o_x, o_y, o_z, kr, a);
            } else {
#pragma unroll
                for (int i = 0; i < kEstRows; ++i) a[i] = 0.f;  // no node: every product 0
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int px = 16 * b + col;
                const float (*f)[64] = feat[wave];
                const float b1 = f[kr][px], b2 = f[4 + kr][px], b4 = f[8 + kr][px], b5 = f[12 + kr][px];
                const float p2 = f[16][px];
                mf4 K = zero;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const mf4 cs = mfma4(a[1 + 2 * j], b2, zero);
                    const mf4 dd = mfma4(a[2 + 2 * j], b1, zero);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float rr = __builtin_amdgcn_rsqf(fmaxf(dd[r] + p2, 1e-24f));
                        K[r] = fmaf(fmaxf(cs[r], 0.f), rr * rr * rr, K[r]);
                    }
                }
                const mf4 dc = mfma4(a[0], b1, zero);
                mf4 inv, i2;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    inv[r] = __builtin_amdgcn_rsqf(fmaxf(dc[r] + p2, 1e-30f));
                    i2[r] = inv[r] * inv[r];
                }
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    const mf4 n1 = mfma4(a[9 + ch], b1, zero);
                    mf4 n2 = mfma4(a[12 + 3 * ch], b1, zero);
                    n2 = mfma4(a[13 + 3 * ch], b4, n2);
                    n2 = mfma4(a[14 + 3 * ch], b5, n2);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        tot[b][ch] = fmaf(fmaxf(fmaf(n1[r], inv[r], n2[r] * i2[r]), 0.f), K[r], tot[b][ch]);
                }
                // one block's products at a time: the blocks' results are not
                // held together
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        c = prev;
    }
    // pixel 16 b + col's sums are spread over lanes col, col + 16, col + 32,
    // col + 48; lane l keeps block l / 16, its own pixel
    float out[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float v = tot[b][ch];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kr == b) out[ch] = v;
        }
    if (live) {
        float4 o = out_rad[q];
        o.x += out[0] * P.br;
        o.y += out[1] * P.bg;
        o.z += out[2] * P.bb;
        out_rad[q] = o;
    }
}

// A frame's pool: no chunk taken, no tile's chain, and the descriptor the
// traversal reads (in device memory, so that it is read only where a chunk
// is handed over).
__global__ void k_pool_reset(EstPool p, EstPool* __restrict__ desc, int ntiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntiles) p.head[i] = -1;
    if (i == 0) {
        *p.taken = 0;
        *desc = p;
    }
}

// The launch order of the cut's tiles from their costs in the last frame,
// heaviest first (longest-processing-time order): the tiles' costs spread
// over two orders of magnitude, and a heavy tile started last runs alone at
// the end of the launch.  A bucket sort on eight buckets per octave of cost,
// in one workgroup; the order within a bucket is that of the LDS atomics.
// A tile's result does not depend on the order.
constexpr int kOrderBins = 256;
constexpr int kOrderThreads = 1024;

__device__ inline int order_bin(float c) {
    const int b = (int)(__log2f(fmaxf(c, 1.f)) * 8.f);
    return kOrderBins - 1 - min(max(b, 0), kOrderBins - 1);
}

__global__ __launch_bounds__(kOrderThreads) void k_tile_order(const float* __restrict__ cost, int n,
                                                              int* __restrict__ order) {
    __shared__ int hist[kOrderBins];
    __shared__ int off[kOrderBins];
    for (int i = threadIdx.x; i < kOrderBins; i += kOrderThreads) hist[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kOrderThreads) atomicAdd(&hist[order_bin(cost[i])], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int b = 0; b < kOrderBins; ++b) {
            off[b] = s;
            s += hist[b];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kOrderThreads) order[atomicAdd(&off[order_bin(cost[i])], 1)] = i;
}

}  // namespace

namespace arc {

int cut_moments_build(arc_gather* g, arc_bvh* b) {
    arc_ctx* ctx = g->ctx;
    const int n = b->n;
    if (n < 2) return ARC_OK;
    if (!b->has_cut_records || !g->crec.p) return fail(ctx, ARC_E_STATE, "moment records need the cut tree");
    const int nblocks = (n + kScanBlock - 1) / kScanBlock;
    ARC_TRY(ensure(ctx, g->mom, sizeof(float4) * kMom * (size_t)(n - 1)));
    ARC_TRY(ensure(ctx, g->mom_pre, sizeof(double) * kMomTerms * (size_t)n));
    ARC_TRY(ensure(ctx, g->mom_tot, sizeof(double) * kMomTerms * (size_t)nblocks));
    hipLaunchKernelGGL(k_mom_scan, dim3(nblocks, kGroups), dim3(kScanBlock), 0, ctx->stream, g->crec.as<float4>(),
                       b->nodes.as<arc_bvh_node>(), n,
                       g->mom_pre.as<double>(), g->mom_tot.as<double>());
    hipLaunchKernelGGL(k_mom_offsets, dim3(kMomTerms), dim3(kScanBlock), 0, ctx->stream, g->mom_tot.as<double>(),
                       nblocks);
    hipLaunchKernelGGL(k_mom_nodes, dim3((n - 1 + kScanBlock - 1) / kScanBlock), dim3(kScanBlock), 0, ctx->stream,
                       b->ranges.as<int>(), b->nodes.as<arc_bvh_node>(), n, g->mom_pre.as<double>(),
                       g->mom_tot.as<double>(), g->mom.as<float4>());
    ARC_LAUNCHED(ctx);
    ARC_TRY(ensure(ctx, g->mom_nbox, sizeof(float4) * 2 * (size_t)(n - 1)));
    ARC_TRY(node_normal_boxes(b, g->crec.as<float4>(), g->mom_nbox.as<float4>()));
    g->mom_ready = true;
    return ARC_OK;
}

int cut_moments(arc_gather* g, arc_bvh* b, const CutEBParams& cp, float4* rad, unsigned long long* counters,
                hipEvent_t mark, CutView* defer) {
    arc_ctx* ctx = g->ctx;
    const int n = b->n;
    if (!g->mom_ready || !g->mom.p || g->mom.bytes < sizeof(float4) * kMom * (size_t)(n - 1))
        return fail(ctx, ARC_E_STATE, "moment estimator: node records not built");
    CutEB t{};
    Threshold th{};
    ARC_TRY(prepare_cut(g, b, cp, t, th));
    const CutMom tm{b->nodes.as<arc_bvh_node>(), g->crec.as<float4>(), g->mom.as<float4>(), n};
    const int blocks = ((g->width + 15) / 16) * ((g->height + 15) / 16);
    const int ntiles = 4 * blocks;
    if (defer) {
        *defer = CutView{tm.nodes, tm.crec, tm.mom, g->mom_nbox.as<float4>(), g->gpos, g->gnrm, g->galb,
                         g->mats.as<arc_material>(),
                         th.est, th.sum, th.eps, th.floor_frac, th.inv_npix, g->width, g->height, cp.delta,
                         cp.budget, cp.criterion, cp.opening * cp.opening, cp.exact_count, nullptr, rad, counters,
                         0, ntiles};
        return ARC_OK;
    }
    if (g->tile_cost_n != ntiles || g->tile_cost.bytes < sizeof(float) * (size_t)ntiles) g->tile_cost_n = 0;
    ARC_TRY(ensure(ctx, g->tile_cost, sizeof(float) * (size_t)ntiles));
    ARC_TRY(ensure(ctx, g->tile_order, sizeof(int) * (size_t)ntiles));
    // the last frame's tile costs, if they describe this image, order the launch
    const int* order = nullptr;
    if (g->tile_cost_n == ntiles) {
        hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, ctx->stream, g->tile_cost.as<float>(), ntiles,
                           g->tile_order.as<int>());
        order = g->tile_order.as<int>();
    }
    // the kept nodes' pool for the matrix-core pass
    const size_t cap = (size_t)ntiles * (size_t)std::max(g->est_chunks, 0);
    EstPool pool{};
    EstPool* desc = nullptr;
    if (cap > 0) {
        if (cap > (size_t)INT_MAX) return fail(ctx, ARC_E_INVALID, "estimate pool too large");
        ARC_TRY(ensure(ctx, g->est_entries, sizeof(uint32_t) * kKeep * cap));
        // + the chunk counter and the descriptor
        ARC_TRY(ensure(ctx, g->est_meta, sizeof(int2) * (cap + 2) + sizeof(EstPool)));
        ARC_TRY(ensure(ctx, g->est_head, sizeof(int) * (size_t)ntiles));
        ARC_TRY(ensure(ctx, g->est_queued, sizeof(float) * (size_t)ntiles));
        ARC_TRY(ensure(ctx, g->est_order, sizeof(int) * (size_t)ntiles));
        int2* meta = g->est_meta.as<int2>();
        pool = EstPool{g->est_entries.as<uint32_t>(), meta, reinterpret_cast<int*>(meta + cap), (int)cap,
                       g->est_head.as<int>(), g->est_queued.as<float>()};
        desc = reinterpret_cast<EstPool*>(meta + cap + 2);
        hipLaunchKernelGGL(k_pool_reset, dim3((ntiles + 255) / 256), dim3(256), 0, ctx->stream, pool, desc, ntiles);
    }
    if (mark) ARC_HIP(ctx, hipEventRecord(mark, ctx->stream));
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kCutThreads), 0, ctx->stream, tm.nodes, tm.crec, tm.mom,
                           g->mom_nbox.as<float4>(), g->gpos,
                           g->gnrm, g->galb, g->mats.as<arc_material>(), g->width, g->height, th, cp.delta, cp.budget,
                           cp.criterion, cp.opening * cp.opening, cp.exact_count, order, g->tile_cost.as<float>(), rad,
                           counters, desc);
    };
    // the traversal's stack depth (its deepest cut): a launch-shape setting
    switch ((cap > 0 ? 1000 : 0) + g->cut_stack) {
        case 64: launch(k_cut_mom_tile<false, 64>); break;
        case 128: launch(k_cut_mom_tile<false, 128>); break;
        case 1064: launch(k_cut_mom_tile<true, 64>); break;
        case 1096: launch(k_cut_mom_tile<true, 96>); break;
        case 1128: launch(k_cut_mom_tile<true, 128>); break;
        default: launch(k_cut_mom_tile<false, 96>); break;
    }
    ARC_LAUNCHED(ctx);
    g->tile_cost_n = ntiles;
    if (cap > 0) {
        // this frame's pooled counts order the estimate pass, heaviest tile first
        hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, ctx->stream, pool.queued, ntiles,
                           g->est_order.as<int>());
        hipLaunchKernelGGL(k_mom_est, dim3(blocks), dim3(kCutThreads), 0, ctx->stream, tm.mom, pool.entries, pool.meta,
                           pool.head, g->gpos, g->gnrm, g->galb, g->mats.as<arc_material>(), g->width, g->height,
                           g->est_order.as<int>(), rad);
        ARC_LAUNCHED(ctx);
    }
    return ARC_OK;
}

int cut_moments_views(arc_ctx* ctx, hipStream_t stream, const CutView* views, int nviews, int nitems,
                      float* costs, int* order, bool ordered) {
    if (nviews < 1 || nitems < 1) return ARC_OK;
    if (ordered) hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, stream, costs, nitems, order);
    hipLaunchKernelGGL(k_cut_mom_views, dim3((nitems + 3) / 4), dim3(kCutThreads), 0, stream, views, nviews,
                       ordered ? order : nullptr, nitems);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

}  // namespace arc
