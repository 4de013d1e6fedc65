// AVPL radiance / antiradiance gather on gfx950 — replaces CudaGather
// (cudaGather.h:25-70, cudaGather.cu:156-1051).
//
// Kernels
//   k_pack        AVPL SoA (AvplsGpuParam) -> 80-byte records: position,
//                 unit normal, raw and unit incident direction, radiance with
//                 the AVPL-side Lambertian BRDF folded in, antiradiance with
//                 the disk normalisation 1/(pi r^2) folded in.
//   k_exact       exact gather (kernel_radiance_simple :269-314 and
//                 kernel_antiradiance :446-557 fused): one pixel per lane,
//                 AVPL records read with wave-uniform (scalar) loads, runs of
//                 64 pairs summed in FP32 and added to the pixel total with
//                 Kahan compensation; optional split over AVPL ranges.
//   k_exact_f64   the same sums in FP64 with the reference's formulas (the
//                 reference image for error metrics).
//   k_finalize    sums the splits (Kahan), writes radiance, antiradiance and
//                 result = radiance - antiradiance (kernel_combine :782-803).
//   k_cut_sa      light-cut radiance with the solid-angle criterion
//                 (kernel_bvh :316-444): per-pixel DFS, stack in LDS.
#include <cmath>
#include <cstring>
#include <vector>

#include "cutnet.hpp"
#include "gather_common.hpp"

using arc::DevBuf;
using arc::f3;
using namespace arc::gm;
using namespace arc::gd;

namespace {

// ------------------------------------------------------------- pack ----
__global__ __launch_bounds__(kThreads) void k_pack(arc::AvplsDev a, const arc_material* __restrict__ mats,
                                                   int general, int model, float inv_pi_r2,
                                                   float4* __restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const f3 p = a.pos[i], q = a.nrm[i], w = a.dir[i], L = a.rad[i];
    const f3 A = a.anti ? a.anti[i] : f3{0.f, 0.f, 0.f};
    const int mi = a.mat[i];
    const float nq = 1.0f / sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
    const float w2 = w.x * w.x + w.y * w.y + w.z * w.z;
    const float nw = w2 > 0.f ? 1.0f / sqrtf(w2) : 0.f;
    float lr = L.x, lg = L.y, lb = L.z;
    if (!general && w2 != 0.f) {  // Lambertian AVPL: f_r = diffuse / pi, constant per AVPL
        const arc_material& m = mats[mat_slot(mi)];
        lr *= kOneOverPi * m.diffuse.x;
        lg *= kOneOverPi * m.diffuse.y;
        lb *= kOneOverPi * m.diffuse.z;
    }
    const float as = model == ARC_ANTIRAD_DISK ? inv_pi_r2 : 1.0f;
    float4* r = rec + (size_t)kRecF4 * i;
    r[0] = make_float4(p.x, p.y, p.z, w.x * nw);
    r[1] = make_float4(q.x * nq, q.y * nq, q.z * nq, w.y * nw);
    r[2] = make_float4(w.x, w.y, w.z, w.z * nw);
    r[3] = make_float4(lr, lg, lb, (float)mi);
    r[4] = make_float4(A.x * as, A.y * as, A.z * as, kGuard * sqrtf(w2));  // guard for |w . n|
}

// ------------------------------------------------------------ exact ----
// part layout: [split][6][npix] = radiance rgb, antiradiance rgb.
template <int MODEL, bool GENERAL, bool DO_RAD, bool DO_ANTI>
__global__ __launch_bounds__(kThreads) void k_exact(const float4* __restrict__ rec, arc::AvplsDev raw, int chunk,
                                                    const float4* __restrict__ gpos,
                                                    const float4* __restrict__ gnrm,
                                                    const float4* __restrict__ galb,
                                                    const arc_material* __restrict__ mats, int npix,
                                                    float camx, float camy, float camz, AntiParams ap,
                                                    float* __restrict__ part) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    const bool live = pix < npix;
    const int q = live ? pix : 0;
    const PixelF P = load_pixel(gpos, gnrm, galb, mats, q);
    const int n_avpl = raw.n;
    const int a0 = blockIdx.y * chunk;
    const int a1 = min(n_avpl, a0 + chunk);
    Kahan kr[3], ka[3];
    for (int base = a0; base < a1; base += kRun) {
        const int end = min(a1, base + kRun);
        float sr0 = 0.f, sr1 = 0.f, sr2 = 0.f, sa0 = 0.f, sa1 = 0.f, sa2 = 0.f;
#pragma unroll 2
        for (int i = base; i < end; ++i) {
            const float4* r = rec + (size_t)kRecF4 * i;
            const float4 r0 = r[0], r1 = r[1], r2v = r[2], r3 = r[3], r4 = r[4];
            // d = a - p, from pixel to AVPL
            const float dx = r0.x - P.px, dy = r0.y - P.py, dz = r0.z - P.pz;
            const float dd = dx * dx + dy * dy + dz * dz;
            const float inv = rsq(dd);
            const float cn = P.nx * dx + P.ny * dy + P.nz * dz;        // |d| cos at the pixel
            const float ca = -(r1.x * dx + r1.y * dy + r1.z * dz);     // |d| cos at the AVPL
            // antiradiance decision and weight
            const float m =
                DO_ANTI ? anti_weight<MODEL>(P, r0, r1, r2v, r4.w, dx, dy, dz, dd, inv, cn, ca, ap, gpos, gnrm, q, raw, i)
                        : 0.f;
            if (GENERAL) {
                using V3 = V<float>;
                const V3 pos = mk(P.px, P.py, P.pz), nrm = mk(P.nx, P.ny, P.nz);
                const V3 apos = mk(r0.x, r0.y, r0.z), anrm = mk(r1.x, r1.y, r1.z);
                const V3 aw = mk(r2v.x, r2v.y, r2v.z);
                const V3 cam = mk(camx, camy, camz);
                Mat<float> pm = material<float>(mats, P.mat);
                if (galb) pm.diffuse = mk(P.br / kOneOverPi, P.bg / kOneOverPi, P.bb / kOneOverPi);
                if (DO_RAD) {
                    const Mat<float> am = material<float>(mats, (int)r3.w);
                    const V3 v = radiance(pos, nrm, pm, apos, anrm, aw, mk(r3.x, r3.y, r3.z), am, cam);
                    sr0 += v.x; sr1 += v.y; sr2 += v.z;
                }
                if (DO_ANTI && m != 0.f) {
                    // record holds A / (pi r^2) (disk) or A (cone); the pixel BRDF varies per pair
                    const V3 b = f_r3(apos, pos, cam, nrm, pm) * m;
                    sa0 += r4.x * b.x; sa1 += r4.y * b.y; sa2 += r4.z * b.z;
                }
            } else {
                if (DO_RAD) {
                    const float c1 = fminf(fmaxf(cn * inv, 0.f), 1.f);
                    const float c2 = fminf(fmaxf(ca * inv, 0.f), 1.f);
                    const float gg = c1 * c2 * (inv * inv);
                    sr0 += r3.x * gg; sr1 += r3.y * gg; sr2 += r3.z * gg;
                }
                if (DO_ANTI) { sa0 += r4.x * m; sa1 += r4.y * m; sa2 += r4.z * m; }
            }
        }
        if (DO_RAD) { kr[0].add(sr0); kr[1].add(sr1); kr[2].add(sr2); }
        if (DO_ANTI) { ka[0].add(sa0); ka[1].add(sa1); ka[2].add(sa2); }
    }
    if (!live) return;
    float* o = part + (size_t)blockIdx.y * 6 * npix + pix;
    if (!GENERAL) {  // the Lambertian pixel BRDF is constant per pixel
        if (DO_RAD) {
            o[0] = kr[0].s * P.br; o[(size_t)npix] = kr[1].s * P.bg; o[2 * (size_t)npix] = kr[2].s * P.bb;
        }
        if (DO_ANTI) {
            o[3 * (size_t)npix] = ka[0].s * P.br; o[4 * (size_t)npix] = ka[1].s * P.bg;
            o[5 * (size_t)npix] = ka[2].s * P.bb;
        }
    } else {
        if (DO_RAD) { o[0] = kr[0].s; o[(size_t)npix] = kr[1].s; o[2 * (size_t)npix] = kr[2].s; }
        if (DO_ANTI) {
            o[3 * (size_t)npix] = ka[0].s; o[4 * (size_t)npix] = ka[1].s; o[5 * (size_t)npix] = ka[2].s;
        }
    }
}

// Sums the splits and writes the float4 targets (kernel_combine semantics:
// result = radiance - antiradiance including w, i.e. 1 - 1 = 0).
// Radiance comes from rad_splits splits of `part`, or from rad_in when
// rad_splits < 0 (the cut kernel wrote it); antiradiance from anti_splits
// splits (0: none).
__global__ __launch_bounds__(kThreads) void k_finalize(const float* __restrict__ part, int rad_splits,
                                                       int anti_splits, int npix,
                                                       const float4* __restrict__ rad_in,
                                                       float4* __restrict__ out_rad,
                                                       float4* __restrict__ out_anti,
                                                       float4* __restrict__ out_res) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix) return;
    Kahan k[6];
    for (int s = 0; s < rad_splits; ++s)
        for (int c = 0; c < 3; ++c) k[c].add(part[((size_t)s * 6 + c) * npix + pix]);
    for (int s = 0; s < anti_splits; ++s)
        for (int c = 3; c < 6; ++c) k[c].add(part[((size_t)s * 6 + c) * npix + pix]);
    const float4 rad = rad_splits >= 0 ? make_float4(k[0].s, k[1].s, k[2].s, 1.f) : rad_in[pix];
    const float4 anti = make_float4(k[3].s, k[4].s, k[5].s, 1.f);
    if (rad_splits >= 0 && out_rad) out_rad[pix] = rad;
    if (out_anti) out_anti[pix] = anti;
    if (out_res)
        out_res[pix] = make_float4(rad.x - anti.x, rad.y - anti.y, rad.z - anti.z, rad.w - anti.w);
}

// --------------------------------------------------------- exact f64 ----
template <int MODEL>
__global__ __launch_bounds__(kThreads) void k_exact_f64(arc::AvplsDev a, const float4* __restrict__ gpos,
                                                        const float4* __restrict__ gnrm,
                                                        const float4* __restrict__ galb,
                                                        const arc_material* __restrict__ mats, int npix,
                                                        float camx, float camy, float camz, float radius,
                                                        double cos_half, int chunk,
                                                        double* __restrict__ part) {
    using V3 = V<double>;
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    const bool live = pix < npix;
    const int q = live ? pix : 0;
    const float4 p4 = gpos[q], n4 = gnrm[q];
    const V3 pos = mk<double>(p4.x, p4.y, p4.z), nrm = mk<double>(n4.x, n4.y, n4.z);
    Mat<double> pm = material<double>(mats, (int)n4.w);
    if (galb) pm.diffuse = mk<double>(galb[q].x, galb[q].y, galb[q].z);
    const V3 cam = mk<double>(camx, camy, camz);
    const int a0 = blockIdx.y * chunk, a1 = min(a.n, a0 + chunk);
    V3 R = mk<double>(0, 0, 0), A = mk<double>(0, 0, 0);
    for (int i = a0; i < a1; ++i) {
        const V3 apos = ld3<double>(a.pos[i]), anrm = ld3<double>(a.nrm[i]), aw = ld3<double>(a.dir[i]);
        const Mat<double> am = material<double>(mats, a.mat[i]);
        R = R + radiance(pos, nrm, pm, apos, anrm, aw, ld3<double>(a.rad[i]), am, cam);
        if (a.anti) {
            const V3 Av = ld3<double>(a.anti[i]);
            if (MODEL == ARC_ANTIRAD_DISK)
                A = A + antirad_disk(pos, nrm, pm, apos, anrm, aw, Av, cam, (double)radius);
            else
                A = A + antirad_cone(pos, nrm, pm, apos, aw, Av, cam, cos_half);
        }
    }
    if (!live) return;
    double* o = part + (size_t)blockIdx.y * 6 * npix + pix;
    o[0] = R.x; o[(size_t)npix] = R.y; o[2 * (size_t)npix] = R.z;
    o[3 * (size_t)npix] = A.x; o[4 * (size_t)npix] = A.y; o[5 * (size_t)npix] = A.z;
}

__global__ __launch_bounds__(kThreads) void k_finalize_f64(const double* __restrict__ part, int splits, int npix,
                                                           double* __restrict__ rad3, double* __restrict__ anti3) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < splits; ++k)
        for (int c = 0; c < 6; ++c) s[c] += part[((size_t)k * 6 + c) * npix + pix];
    for (int c = 0; c < 3; ++c) {
        if (rad3) rad3[3 * (size_t)pix + c] = s[c];
        if (anti3) anti3[3 * (size_t)pix + c] = s[3 + c];
    }
}

// ------------------------------------------------------ cut (solid) ----
// Decision arithmetic of kernel_bvh (cudaGather.cu:405-414) in FP32 with IEEE
// division and square root and no fused multiply-add, so every traversal
// decision equals the oracle's REF32 decision.
__device__ __attribute__((noinline)) bool cluster_visible(f3 n, f3 p, f3 bmin, f3 bmax) {
#pragma clang fp contract(off)
    const float ax = bmax.x - p.x, ay = bmax.y - p.y, az = bmax.z - p.z;
    const float sa = 1.0f / sqrtf(ax * ax + ay * ay + az * az);
    const float c1 = n.x * (ax * sa) + n.y * (ay * sa) + n.z * (az * sa);
    const float bx = bmin.x - p.x, by = bmin.y - p.y, bz = bmin.z - p.z;
    const float sb = 1.0f / sqrtf(bx * bx + by * by + bz * bz);
    const float c2 = n.x * (bx * sb) + n.y * (by * sb) + n.z * (bz * sb);
    return (c1 > 0.f) || (c2 > 0.f);
}

__device__ __attribute__((noinline)) bool solid_angle_below(f3 p, f3 rep, f3 bmin, f3 bmax, float thresh) {
#pragma clang fp contract(off)
    const float ex = bmax.x - bmin.x, ey = bmax.y - bmin.y, ez = bmax.z - bmin.z;
    const float radius = 0.5f * sqrtf(ex * ex + ey * ey + ez * ez);
    const float dx = rep.x - p.x, dy = rep.y - p.y, dz = rep.z - p.z;
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    const float omega = (2.0f * kPi) * (1.0f - dist / sqrtf(radius * radius + dist * dist));
    return omega < thresh;
}

struct CutTree {
    const arc_bvh_node* nodes;
    const int* mat;
    const f3* pos;
    const f3* nrm;
    const f3* rad;
    const f3* dir;
    int n;  // leaves
};

__global__ __launch_bounds__(kThreads) void k_cut_sa(CutTree t, const float4* __restrict__ gpos,
                                                     const float4* __restrict__ gnrm,
                                                     const float4* __restrict__ galb,
                                                     const arc_material* __restrict__ mats, int W, int H,
                                                     float camx, float camy, float camz, float thresh,
                                                     int stack_size, int sentinel,
                                                     float4* __restrict__ out_rad,
                                                     unsigned long long* __restrict__ counters) {
    extern __shared__ int stk[];  // [stack_size][kThreads]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool live = x < W && y < H;
    unsigned long long nvis = 0, neval = 0, nerr = 0;
    if (live) {
        using V3 = V<float>;
        const int pix = y * W + x;
        const float4 p4 = gpos[pix], n4 = gnrm[pix];
        const f3 p = f3of(p4), nraw = f3of(n4);
        const V3 pos = mk(p4.x, p4.y, p4.z), nrm = mk(n4.x, n4.y, n4.z);
        Mat<float> pm = material<float>(mats, (int)n4.w);
        if (galb) pm.diffuse = mk(galb[pix].x, galb[pix].y, galb[pix].z);
        const V3 cam = mk(camx, camy, camz);
        auto eval = [&](int idx) {
            const Mat<float> am = material<float>(mats, t.mat[idx]);
            return radiance(pos, nrm, pm, ld3<float>(t.pos[idx]), ld3<float>(t.nrm[idx]), ld3<float>(t.dir[idx]),
                            ld3<float>(t.rad[idx]), am, cam);
        };
        V3 R = mk(0.f, 0.f, 0.f);
        int sp = 0;
        stk[sp * kThreads + tid] = -1;
        sp = 1;
        bool error = false;
        while (sp > 0) {
            const int node = stk[(--sp) * kThreads + tid];
            if (node >= 0) {
                R = R + eval(node);
                neval++;
                continue;
            }
            const int k = -(node + 1);
            if (k >= t.n - 1) {
                error = true;
                break;
            }
            nvis++;
            const arc_bvh_node b = t.nodes[k];
            const f3 bmin = f3{b.bbMin.x, b.bbMin.y, b.bbMin.z}, bmax = f3{b.bbMax.x, b.bbMax.y, b.bbMax.z};
            if (!cluster_visible(nraw, p, bmin, bmax)) continue;
            const int idx = t.n + k;
            if (solid_angle_below(p, t.pos[idx], bmin, bmax, thresh)) {
                R = R + eval(idx);
                neval++;
            } else {
                stk[sp * kThreads + tid] = b.left;
                stk[(sp + 1) * kThreads + tid] = b.right;
                sp += 2;
                if (sp >= stack_size) {
                    error = true;
                    break;
                }
            }
        }
        float4 o = make_float4(R.x, R.y, R.z, 1.f);
        if (error) {
            nerr = 1;
            if (sentinel) o = make_float4(100000.f, 0.f, 100000.f, 1.f);
        }
        out_rad[pix] = o;
    }
    if (counters) {
        // one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            nvis += __shfl_down(nvis, off);
            neval += __shfl_down(neval, off);
            nerr += __shfl_down(nerr, off);
        }
        if (lane == 0) {
            atomicAdd(&counters[C_VISITED], nvis);
            atomicAdd(&counters[C_EVALUATED], neval);
            atomicAdd(&counters[C_ERRORS], nerr);
        }
    }
}

// ------------------------------------------------------------- host ----
inline int blocks(long long n) { return (int)((n + kThreads - 1) / kThreads); }

int pick_splits(int npix, int n_avpl) {
    const int pblocks = blocks(npix);
    int s = (2048 + pblocks - 1) / pblocks;
    s = std::max(1, std::min(s, (n_avpl + 255) / 256));
    return s;
}

template <int MODEL, bool GENERAL, bool DO_RAD, bool DO_ANTI>
void launch_exact(arc_gather* g, const arc::AvplsDev& a, int chunk, int splits, float3 cam, float radius,
                  double cos_half) {
    const int npix = g->width * g->height;
    hipLaunchKernelGGL((k_exact<MODEL, GENERAL, DO_RAD, DO_ANTI>), dim3(blocks(npix), splits), dim3(kThreads), 0,
                       g->ctx->stream, g->rec.as<float4>(), a, chunk, g->gpos, g->gnrm, g->galb,
                       g->mats.as<arc_material>(), npix, cam.x, cam.y, cam.z, anti_params(radius, cos_half),
                       g->part.as<float>());
}

template <bool DO_RAD, bool DO_ANTI>
void dispatch_exact(arc_gather* g, int model, const arc::AvplsDev& a, int chunk, int splits, float3 cam,
                    float radius, double cos_half) {
    if (model == ARC_ANTIRAD_DISK) {
        if (g->general) launch_exact<ARC_ANTIRAD_DISK, true, DO_RAD, DO_ANTI>(g, a, chunk, splits, cam, radius, cos_half);
        else launch_exact<ARC_ANTIRAD_DISK, false, DO_RAD, DO_ANTI>(g, a, chunk, splits, cam, radius, cos_half);
    } else {
        if (g->general) launch_exact<ARC_ANTIRAD_CONE, true, DO_RAD, DO_ANTI>(g, a, chunk, splits, cam, radius, cos_half);
        else launch_exact<ARC_ANTIRAD_CONE, false, DO_RAD, DO_ANTI>(g, a, chunk, splits, cam, radius, cos_half);
    }
}

// This frame's LBVH for the error-bound and learned cuts, with their node
// records in g->crec.
int build_cut_tree(arc_gather* g, const arc::AvplsDev& a, bool validate) {
    g->mom_ready = false;
    if (a.n > 1) ARC_TRY(arc::ensure(g->ctx, g->crec, sizeof(float4) * arc::kCrec * (size_t)(2 * a.n - 1)));
    return arc::bvh_build_cut(g->bvh, a, validate, g->mats.as<arc_material>(), g->crec.as<float4>());
}

// AVPL pointers on the device, staging host input when asked (AvplsGpu,
// data_gpu.cu:5-52).
int stage_avpls(arc_gather* g, const arc_avpls* in, int flags, arc::AvplsDev& out) {
    arc_ctx* ctx = g->ctx;
    const int n = in->numAvpls;
    out.n = n;
    if (!(flags & ARC_INPUT_HOST)) {
        out.pos = reinterpret_cast<const f3*>(in->position);
        out.nrm = reinterpret_cast<const f3*>(in->normal);
        out.dir = reinterpret_cast<const f3*>(in->incDirection);
        out.rad = reinterpret_cast<const f3*>(in->incRadiance);
        out.anti = reinterpret_cast<const f3*>(in->antiradiance);
        out.mat = in->materialIndex;
        out.bounce = in->bounce;
        return ARC_OK;
    }
    const size_t v3 = sizeof(f3) * (size_t)n, vi = sizeof(int) * (size_t)n;
    ARC_TRY(arc::ensure(ctx, g->stage, 5 * v3 + 2 * vi));
    char* base = g->stage.as<char>();
    const void* src[7] = {in->position, in->normal, in->incDirection, in->incRadiance, in->antiradiance,
                          in->materialIndex, in->bounce};
    const size_t sz[7] = {v3, v3, v3, v3, v3, vi, vi};
    char* dst[7];
    size_t off = 0;
    for (int k = 0; k < 7; ++k) {
        dst[k] = base + off;
        off += sz[k];
        if (src[k]) ARC_HIP(ctx, hipMemcpyAsync(dst[k], src[k], sz[k], hipMemcpyHostToDevice, ctx->stream));
    }
    // the caller may free or reuse its host arrays once the call returns
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out.pos = reinterpret_cast<const f3*>(dst[0]);
    out.nrm = reinterpret_cast<const f3*>(dst[1]);
    out.dir = reinterpret_cast<const f3*>(dst[2]);
    out.rad = reinterpret_cast<const f3*>(dst[3]);
    out.anti = in->antiradiance ? reinterpret_cast<const f3*>(dst[4]) : nullptr;
    out.mat = reinterpret_cast<const int*>(dst[5]);
    out.bounce = in->bounce ? reinterpret_cast<const int*>(dst[6]) : nullptr;
    return ARC_OK;
}

int validate_params(arc_gather* g, const arc_avpls* avpls, const arc_gather_params* p) {
    if (!g || !avpls || !p) return ARC_E_INVALID;
    if (!g->gpos || !g->gnrm) return arc::fail(g->ctx, ARC_E_STATE, "no G-buffer bound");
    if (avpls->numAvpls < 0) return arc::fail(g->ctx, ARC_E_INVALID, "numAvpls < 0");
    if (avpls->numAvpls > 0 && (!avpls->position || !avpls->normal || !avpls->incDirection ||
                                !avpls->incRadiance || !avpls->materialIndex))
        return arc::fail(g->ctx, ARC_E_INVALID, "AVPL arrays missing");
    if (p->antiradianceModel != ARC_ANTIRAD_DISK && p->antiradianceModel != ARC_ANTIRAD_CONE)
        return arc::fail(g->ctx, ARC_E_INVALID, "antiradianceModel");
    if (p->stackSize < 2 || p->stackSize > 64) return arc::fail(g->ctx, ARC_E_INVALID, "stackSize out of [2, 64]");
    if (p->antiradianceMode != ARC_ANTI_EXACT && p->antiradianceMode != ARC_ANTI_CLUSTERED)
        return arc::fail(g->ctx, ARC_E_INVALID, "antiradianceMode");
    return ARC_OK;
}

float4* target(DevBuf& internal, float4* user, arc_gather* g, int& status) {
    if (user) return user;
    status = arc::ensure(g->ctx, internal, sizeof(float4) * (size_t)g->width * g->height);
    return internal.as<float4>();
}

}  // namespace

extern "C" {

void arc_gather_params_default(arc_gather_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->useLightCuts = 1;                        // CConfigManager.cpp:15
    p->cutMode = ARC_CUT_SOLID_ANGLE;
    p->clusterRefinementThreshold = (float)(M_PI / 10.0);  // CConfigManager.cpp:37
    p->errorBound = 0.01f;
    p->antiradianceModel = ARC_ANTIRAD_DISK;
    p->antiradianceMode = ARC_ANTI_CLUSTERED;   // useClusteredDeferred = 1, CConfigManager.cpp:16
    p->photonRadiusScale = 2.0f;                // CConfigManager.cpp:26
    p->coneHalfAngle = (float)(M_PI / 18.0);
    p->writeSentinel = 1;
    p->stackSize = 32;                          // STACK_SIZE, cudaGather.cu:19
    p->seed = 5489;
    p->cutEstimateDepth = 6;
    p->cutBudget = 1 << 16;
    p->cutPerTile = 0;
    p->cutCriterion = ARC_CRITERION_ESTIMATE;
    p->cutImageFloor = 1.0f;
    p->cutConfidence = 0.9f;
    p->cutEvalDepth = 0;
    p->cutEstimator = ARC_ESTIMATOR_CENTROID;
    p->cutExactCount = 16;
    p->cutOpening = 0.3f;
}

int arc_gather_create(arc_ctx* ctx, int width, int height, const arc_material* materials, int num_materials,
                      arc_gather** out) {
    if (!ctx || !out || width <= 0 || height <= 0 || !materials || num_materials <= 0 || num_materials > 256)
        return ARC_E_INVALID;
    arc_gather* g = new arc_gather();
    g->ctx = ctx;
    g->width = width;
    g->height = height;
    g->num_mats = num_materials;
    for (int i = 0; i < num_materials; ++i) {
        const arc_float3& s = materials[i].specular;
        if (std::sqrt(s.x * s.x + s.y * s.y + s.z * s.z) > 0.f) g->general = true;
    }
    // zero-padded to kMaxMaterials entries: kernels mask indices into the table
    std::vector<arc_material> table(kMaxMaterials);
    std::memset(table.data(), 0, sizeof(arc_material) * table.size());
    std::memcpy(table.data(), materials, sizeof(arc_material) * (size_t)num_materials);
    int st = arc::ensure(ctx, g->mats, sizeof(arc_material) * table.size());
    if (st == ARC_OK) {
        hipError_t e = hipMemcpy(g->mats.p, table.data(), sizeof(arc_material) * table.size(),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) st = arc::hip_fail(ctx, e, "hipMemcpy(materials)", __FILE__, __LINE__);
    }
    if (st == ARC_OK) st = arc::ensure(ctx, g->counters, sizeof(unsigned long long) * C_NUM);
    for (int f = 0; st == ARC_OK && f < arc_gather::kRing; ++f)
        for (int i = 0; st == ARC_OK && i < arc_gather::kMarks; ++i) {
            hipError_t e = hipEventCreate(&g->ev[f][i]);
            if (e != hipSuccess) st = arc::hip_fail(ctx, e, "hipEventCreate", __FILE__, __LINE__);
        }
    if (st == ARC_OK) st = arc_bvh_create(ctx, 0, &g->bvh);
    // the frame's tree is rebuilt beside the antiradiance on the second stream
    if (st == ARC_OK) g->bvh->refit_top_threads = 256;
    if (st == ARC_OK) {
        hipError_t e = hipStreamCreateWithFlags(&g->aux, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->join, hipEventDisableTiming);
        if (e != hipSuccess) st = arc::hip_fail(ctx, e, "aux stream", __FILE__, __LINE__);
    }
    if (st != ARC_OK) {
        arc_gather_destroy(g);
        return st;
    }
    *out = g;
    return ARC_OK;
}

void arc_gather_destroy(arc_gather* g) {
    if (!g) return;
    if (g->aux) {
        (void)hipStreamSynchronize(g->aux);
        (void)hipStreamDestroy(g->aux);
    }
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->join) (void)hipEventDestroy(g->join);
    for (int f = 0; f < arc_gather::kRing; ++f)
        for (int i = 0; i < arc_gather::kMarks; ++i)
            if (g->ev[f][i]) (void)hipEventDestroy(g->ev[f][i]);
    arc::release(g->mats);
    arc::release(g->int_rad);
    arc::release(g->int_anti);
    arc::release(g->int_res);
    arc::release(g->rec);
    arc::release(g->part);
    arc::release(g->counters);
    arc::release(g->stage);
    arc::release(g->anti_idx);
    arc::release(g->cbox);
    arc::release(g->ccount);
    arc::release(g->clist);
    arc::release(g->crec);
    arc::release(g->mom);
    arc::release(g->mom_pre);
    arc::release(g->mom_tot);
    arc::release(g->mom_nbox);
    arc::release(g->tile_cost);
    arc::release(g->tile_order);
    arc::release(g->est_entries);
    arc::release(g->est_meta);
    arc::release(g->est_head);
    arc::release(g->est_queued);
    arc::release(g->est_order);
    arc::release(g->cest);
    arc_bvh_destroy(g->bvh);
    delete g;
}

int arc_gather_set_gbuffer(arc_gather* g, const float* pos4, const float* nrm4, const float* albedo4) {
    if (!g || !pos4 || !nrm4) return ARC_E_INVALID;
    g->gpos = reinterpret_cast<const float4*>(pos4);
    g->gnrm = reinterpret_cast<const float4*>(nrm4);
    g->galb = reinterpret_cast<const float4*>(albedo4);
    return ARC_OK;
}

int arc_gather_set_outputs(arc_gather* g, float* result4, float* radiance4, float* antiradiance4) {
    if (!g) return ARC_E_INVALID;
    g->out_res = reinterpret_cast<float4*>(result4);
    g->out_rad = reinterpret_cast<float4*>(radiance4);
    g->out_anti = reinterpret_cast<float4*>(antiradiance4);
    return ARC_OK;
}

arc_bvh* arc_gather_get_bvh(arc_gather* g) { return (g && g->bvh && g->bvh->n > 0) ? g->bvh : nullptr; }

int arc_gather_copy_cut_records(arc_gather* g, int which, float* host, long long max_float4, long long* count) {
    if (!g || !count || max_float4 < 0 || (max_float4 > 0 && !host)) return ARC_E_INVALID;
    if (which != ARC_RECORDS_CENTROID && which != ARC_RECORDS_MOMENTS) return ARC_E_INVALID;
    arc_ctx* ctx = g->ctx;
    const int n = g->bvh ? g->bvh->n : 0;
    *count = 0;
    if (n < 2 || !g->bvh->has_cut_records) return arc::fail(ctx, ARC_E_STATE, "no error-bound cut frame has run");
    const arc::DevBuf& buf = which == ARC_RECORDS_CENTROID ? g->crec : g->mom;
    const long long avail = which == ARC_RECORDS_CENTROID ? (long long)arc::kCrec * (2LL * n - 1) : 13LL * (n - 1);
    if (!buf.p || buf.bytes < sizeof(float4) * (size_t)avail || (which == ARC_RECORDS_MOMENTS && !g->mom_ready))
        return arc::fail(ctx, ARC_E_STATE, "records of this kind were not built by the last frame");
    *count = avail;
    const long long m = std::min(avail, max_float4);
    if (m > 0) {
        ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ARC_HIP(ctx, hipMemcpy(host, buf.p, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost));
    }
    return ARC_OK;
}

int arc_gather_run(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                   const arc_gather_params* p, int flags, arc_gather_stats* stats) {
    ARC_TRY(validate_params(g, avpls, p));
    arc_ctx* ctx = g->ctx;
    const bool profile = (flags & (ARC_RUN_PROFILE | ARC_RUN_TIMELINE)) != 0;
    const int slot = g->ring_next;
    hipEvent_t* ev = g->ev[slot];
    const bool count = (flags & ARC_RUN_COUNTERS) != 0;
    const int npix = g->width * g->height;
    const int n = avpls->numAvpls;
    const float radius = p->photonRadiusScale * sceneExtent / 100.f;  // cudaGather.cu:949
    const double cos_half = std::cos((double)p->coneHalfAngle);
    const float3 cam = make_float3(camPos.x, camPos.y, camPos.z);
    int st = ARC_OK;
    float4* rad = target(g->int_rad, g->out_rad, g, st);
    ARC_TRY(st);
    float4* res = g->out_res;
    float4* anti = g->out_anti;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (count) ARC_HIP(ctx, hipMemsetAsync(g->counters.p, 0, sizeof(unsigned long long) * C_NUM, ctx->stream));
    if (profile) ARC_HIP(ctx, hipEventRecord(ev[0], ctx->stream));

    arc::AvplsDev a{};
    ARC_TRY(stage_avpls(g, avpls, flags, a));
    if (n > 0) {
        ARC_TRY(arc::ensure(ctx, g->rec, sizeof(float4) * kRecF4 * (size_t)n));
        hipLaunchKernelGGL(k_pack, dim3(blocks(n)), dim3(kThreads), 0, ctx->stream, a, g->mats.as<arc_material>(),
                           g->general ? 1 : 0, p->antiradianceModel, 1.0f / (kPi * radius * radius),
                           g->rec.as<float4>());
        ARC_LAUNCHED(ctx);
    }
    if (profile) ARC_HIP(ctx, hipEventRecord(ev[1], ctx->stream));

    const int splits = pick_splits(npix, std::max(n, 1));
    const int chunk = (std::max(n, 1) + splits - 1) / splits;
    ARC_TRY(arc::ensure(ctx, g->part, sizeof(float) * 6 * (size_t)splits * npix));
    unsigned long long* counters = count ? g->counters.as<unsigned long long>() : nullptr;

    const bool cut = p->useLightCuts && n > 1;
    const bool clustered = p->antiradianceMode == ARC_ANTI_CLUSTERED && n > 0;
    // the clustered antiradiance of a cut frame needs neither the LBVH nor the
    // radiance: it runs on the gather's second stream meanwhile.  Not while
    // the stream is being captured into a graph (arc_batch_capture): there
    // the frame stays on one stream.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    ARC_HIP(ctx, hipStreamIsCapturing(ctx->stream, &capture));
    const bool overlap = cut && clustered && capture == hipStreamCaptureStatusNone;
    if (p->useLightCuts && p->cutMode != ARC_CUT_SOLID_ANGLE && p->cutMode != ARC_CUT_ERROR_BOUND &&
        p->cutMode != ARC_CUT_LEARNED)
        return arc::fail(ctx, ARC_E_INVALID, "cutMode");
    const bool moments = p->cutMode != ARC_CUT_SOLID_ANGLE;
    if (p->useLightCuts && moments && g->general)
        return arc::fail(ctx, ARC_E_UNSUPPORTED, "the error-bound and learned cuts need Lambertian materials");
    if (p->useLightCuts && p->cutMode == ARC_CUT_LEARNED && !(g->net && g->net->ready))
        return arc::fail(ctx, ARC_E_STATE, "learned cut without a trained arc_cutnet (arc_gather_set_cutnet)");
    if (p->useLightCuts && p->cutMode == ARC_CUT_LEARNED && !(p->cutConfidence >= 0.5f && p->cutConfidence <= 1.f))
        return arc::fail(ctx, ARC_E_INVALID, "cutConfidence");
    if (p->cutEstimateDepth < 0 || p->cutEstimateDepth > 40) return arc::fail(ctx, ARC_E_INVALID, "cutEstimateDepth");
    if (!(p->cutImageFloor >= 0.f)) return arc::fail(ctx, ARC_E_INVALID, "cutImageFloor");
    if (p->cutEvalDepth < 0 || p->cutEvalDepth > 4) return arc::fail(ctx, ARC_E_INVALID, "cutEvalDepth");
    if (p->cutEstimator != ARC_ESTIMATOR_CENTROID && p->cutEstimator != ARC_ESTIMATOR_MOMENTS)
        return arc::fail(ctx, ARC_E_INVALID, "cutEstimator");
    const bool mom_est = p->useLightCuts && p->cutMode == ARC_CUT_ERROR_BOUND && p->cutEstimator == ARC_ESTIMATOR_MOMENTS;
    if (g->defer_cut && (!mom_est || !cut || profile || stats || count))
        return arc::fail(ctx, ARC_E_INVALID, "a frame whose cut joins a batched launch must be a moment-cut frame "
                                             "without profiling or statistics");
    if (p->cutCriterion != ARC_CRITERION_BOUND && p->cutCriterion != ARC_CRITERION_ESTIMATE &&
        p->cutCriterion != ARC_CRITERION_EXPANSION)
        return arc::fail(ctx, ARC_E_INVALID, "cutCriterion");
    if (mom_est && p->cutCriterion == ARC_CRITERION_ESTIMATE)
        return arc::fail(ctx, ARC_E_INVALID, "the moment estimator takes ARC_CRITERION_BOUND or ARC_CRITERION_EXPANSION");
    if (!mom_est && p->useLightCuts && p->cutMode == ARC_CUT_ERROR_BOUND && p->cutCriterion == ARC_CRITERION_EXPANSION)
        return arc::fail(ctx, ARC_E_INVALID, "ARC_CRITERION_EXPANSION needs the moment estimator");
    if (mom_est && (!p->cutPerTile || p->cutEvalDepth != 0))
        return arc::fail(ctx, ARC_E_INVALID, "the moment estimator runs per 8x8 tile (cutPerTile 1, cutEvalDepth 0)");
    if (p->cutExactCount < 0 || (mom_est && p->cutExactCount > 128))
        return arc::fail(ctx, ARC_E_INVALID, "cutExactCount out of [0, 128]");
    if (!(p->cutOpening > 0.f)) return arc::fail(ctx, ARC_E_INVALID, "cutOpening");
    if (p->cutBudget < 1) return arc::fail(ctx, ARC_E_INVALID, "cutBudget");
    // ---- radiance: light cut over this frame's AvplBvh (cudaGather.cu:854-923) or exact
    int rad_splits = -1, anti_splits = 0;
    if (overlap) {
        ARC_HIP(ctx, hipEventRecord(g->fork, ctx->stream));
        ARC_HIP(ctx, hipStreamWaitEvent(g->aux, g->fork, 0));
        hipStream_t main_stream = ctx->stream;
        ctx->stream = g->aux;  // the antiradiance launches go to the second stream
        const int sa = arc::anti_clustered(g, a, p->antiradianceModel, cam, anti_params(radius, cos_half),
                                           (double)p->coneHalfAngle, sceneExtent, g->part.as<float>(), counters,
                                           profile ? ev[7] : nullptr);
        hipError_t e = hipSuccess;
        if (sa == ARC_OK && profile) e = hipEventRecord(ev[8], g->aux);
        if (sa == ARC_OK && e == hipSuccess) e = hipEventRecord(g->join, g->aux);
        ctx->stream = main_stream;
        ARC_TRY(sa);
        if (e != hipSuccess) return arc::hip_fail(ctx, e, "antiradiance stream", __FILE__, __LINE__);
        anti_splits = 1;
    }
    if (cut) {
        // AvplBvh(avpls, false) per frame (cudaGather.cu:858)
        if (moments)
            ARC_TRY(build_cut_tree(g, a, (flags & ARC_TRUST_INPUT) == 0));
        if (mom_est) ARC_TRY(arc::cut_moments_build(g, g->bvh));
        if (!moments)
            ARC_TRY(arc::bvh_build_avpls(g->bvh, a, nullptr, p->seed, (flags & ARC_TRUST_INPUT) == 0));
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
        arc_bvh* b = g->bvh;
        if (moments) {
            const arc::CutEBParams cp{p->errorBound, p->cutEstimateDepth, 1e-3f * sceneExtent, p->cutBudget,
                                      p->cutPerTile, p->cutCriterion, p->cutImageFloor, p->cutConfidence,
                                      p->cutEvalDepth, p->cutExactCount, p->cutOpening};
            if (p->cutMode == ARC_CUT_LEARNED)
                ARC_TRY(arc::cut_learned(g, b, cp, rad, counters, profile ? ev[6] : nullptr));
            else if (mom_est)
                ARC_TRY(arc::cut_moments(g, b, cp, rad, counters, profile ? ev[6] : nullptr,
                                         g->defer_cut ? &g->def_view : nullptr));
            else
                ARC_TRY(arc::cut_error_bound(g, b, cp, rad, counters, profile ? ev[6] : nullptr));
        } else {
            if (profile) ARC_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
            CutTree t{b->nodes.as<arc_bvh_node>(), b->nd_mat.as<int>(), b->nd_pos.as<f3>(), b->nd_nrm.as<f3>(),
                      b->nd_rad.as<f3>(), b->nd_dir.as<f3>(), b->n};
            dim3 grid((g->width + 15) / 16, (g->height + 15) / 16);
            const size_t lds = sizeof(int) * (size_t)p->stackSize * kThreads;
            hipLaunchKernelGGL(k_cut_sa, grid, dim3(kThreads), lds, ctx->stream, t, g->gpos, g->gnrm, g->galb,
                               g->mats.as<arc_material>(), g->width, g->height, cam.x, cam.y, cam.z,
                               p->clusterRefinementThreshold, p->stackSize, p->writeSentinel, rad, counters);
        }
        ARC_LAUNCHED(ctx);
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
    } else {
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
        rad_splits = n > 0 ? splits : 0;
        if (n > 0 && !clustered) {  // one pass sums both (kernel_radiance_simple + kernel_antiradiance)
            dispatch_exact<true, true>(g, p->antiradianceModel, a, chunk, splits, cam, radius, cos_half);
            anti_splits = splits;
        } else if (n > 0) {
            dispatch_exact<true, false>(g, p->antiradianceModel, a, chunk, splits, cam, radius, cos_half);
        }
        ARC_LAUNCHED(ctx);
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
    }
    // ---- antiradiance: clustered lists (kernel_antiradiance_clusterred) or exact
    if (overlap) {
        ARC_HIP(ctx, hipStreamWaitEvent(ctx->stream, g->join, 0));  // what the radiance did not hide
    } else if (clustered) {
        ARC_TRY(arc::anti_clustered(g, a, p->antiradianceModel, cam, anti_params(radius, cos_half),
                                    (double)p->coneHalfAngle, sceneExtent, g->part.as<float>(), counters,
                                    profile ? ev[7] : nullptr));
        anti_splits = 1;
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[8], ctx->stream));
    } else if (cut && n > 0) {
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[7], ctx->stream));
        dispatch_exact<false, true>(g, p->antiradianceModel, a, chunk, splits, cam, radius, cos_half);
        anti_splits = splits;
        ARC_LAUNCHED(ctx);
        if (profile) ARC_HIP(ctx, hipEventRecord(ev[8], ctx->stream));
    } else if (profile) {
        ARC_HIP(ctx, hipEventRecord(ev[7], ctx->stream));  // summed with the radiance: no pass of its own
        ARC_HIP(ctx, hipEventRecord(ev[8], ctx->stream));
    }
    if (g->defer_cut) {
        // the cut of this frame is launched with other views' (arc_batch);
        // arc::gather_finish combines once it has run
        g->def_rad_splits = rad_splits;
        g->def_anti_splits = anti_splits;
        g->def_rad = rad;
        g->deferred = true;
        return ARC_OK;
    }
    if (profile) ARC_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
    hipLaunchKernelGGL(k_finalize, dim3(blocks(npix)), dim3(kThreads), 0, ctx->stream, g->part.as<float>(),
                       rad_splits, anti_splits, npix, rad, rad, anti, res);
    ARC_LAUNCHED(ctx);
    if (profile) ARC_HIP(ctx, hipEventRecord(ev[5], ctx->stream));
    if (stats) {
        stats->numLeafs = cut ? g->bvh->n : 0;
        if (count) {
            unsigned long long c[C_NUM];
            ARC_HIP(ctx, hipMemcpyAsync(c, g->counters.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
            ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            stats->visitedNodes = (long long)c[C_VISITED];
            stats->evaluatedNodes = (long long)c[C_EVALUATED];
            stats->antiVisitedNodes = (long long)c[C_ANTI_VISITED];
            stats->antiEvaluatedNodes = (long long)c[C_ANTI_EVALUATED];
            stats->errorPixels = (long long)c[C_ERRORS];
            stats->budgetHits = (long long)c[C_BUDGET];
            stats->fallbackNodes = (long long)c[C_FALLBACK];
            stats->testedNodes = (long long)c[C_TESTED];
            stats->exactLeaves = (long long)c[C_EXACT];
        }
        if (flags & ARC_RUN_PROFILE) {
            ARC_HIP(ctx, hipEventSynchronize(ev[5]));
            float t[5];
            for (int i = 0; i < 5; ++i) ARC_HIP(ctx, hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
            stats->msUpload = t[0];
            stats->msBvh = t[1];
            stats->msRadiance = t[2];
            stats->msAntiradiance = t[3];
            stats->msCombine = t[4];
            ARC_HIP(ctx, hipEventElapsedTime(&stats->msTotal, ev[0], ev[5]));
        }
    }
    if (profile) {
        g->ring_next = (slot + 1) % arc_gather::kRing;
        g->ring_count = std::min(g->ring_count + 1, arc_gather::kRing);
    }
    return ARC_OK;
}

}  // extern "C"

namespace arc {

int gather_finish(arc_gather* g) {
    if (!g || !g->deferred) return ARC_E_STATE;
    arc_ctx* ctx = g->ctx;
    const int npix = g->width * g->height;
    hipLaunchKernelGGL(k_finalize, dim3(blocks(npix)), dim3(kThreads), 0, ctx->stream, g->part.as<float>(),
                       g->def_rad_splits, g->def_anti_splits, npix, g->def_rad, g->def_rad, g->out_anti, g->out_res);
    ARC_LAUNCHED(ctx);
    g->deferred = false;
    return ARC_OK;
}

}  // namespace arc

extern "C" {

int arc_gather_profile_history(arc_gather* g, int frames, float* out) {
    if (!g || frames < 0 || (frames > 0 && !out)) return ARC_E_INVALID;
    arc_ctx* ctx = g->ctx;
    const int n = std::min(frames, g->ring_count);
    for (int k = 0; k < n; ++k) {
        const int slot = (g->ring_next - n + k + arc_gather::kRing) % arc_gather::kRing;
        hipEvent_t* ev = g->ev[slot];
        ARC_HIP(ctx, hipEventSynchronize(ev[5]));
        for (int i = 0; i < 5; ++i) ARC_HIP(ctx, hipEventElapsedTime(&out[6 * k + i], ev[i], ev[i + 1]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[6 * k + 5], ev[0], ev[5]));
    }
    return n;
}

arc_ctx* arc_gather_ctx(arc_gather* g) { return g ? g->ctx : nullptr; }

int arc_gather_kernel_history(arc_gather* g, int frames, float* out) {
    if (!g || frames < 0 || (frames > 0 && !out)) return ARC_E_INVALID;
    arc_ctx* ctx = g->ctx;
    const int n = std::min(frames, g->ring_count);
    for (int k = 0; k < n; ++k) {
        const int slot = (g->ring_next - n + k + arc_gather::kRing) % arc_gather::kRing;
        hipEvent_t* ev = g->ev[slot];
        ARC_HIP(ctx, hipEventSynchronize(ev[5]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[2 * k + 0], ev[6], ev[3]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[2 * k + 1], ev[7], ev[8]));
    }
    return n;
}

int arc_gather_kernel_spans(arc_gather* g, int frames, float* out) {
    if (!g || frames < 0 || (frames > 0 && !out)) return ARC_E_INVALID;
    arc_ctx* ctx = g->ctx;
    const int n = std::min(frames, g->ring_count);
    for (int k = 0; k < n; ++k) {
        const int slot = (g->ring_next - n + k + arc_gather::kRing) % arc_gather::kRing;
        hipEvent_t* ev = g->ev[slot];
        ARC_HIP(ctx, hipEventSynchronize(ev[5]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[4 * k + 0], ev[0], ev[6]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[4 * k + 1], ev[0], ev[3]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[4 * k + 2], ev[0], ev[7]));
        ARC_HIP(ctx, hipEventElapsedTime(&out[4 * k + 3], ev[0], ev[8]));
    }
    return n;
}

int arc_gather_set_cutnet(arc_gather* g, arc_cutnet* net) {
    if (!g) return ARC_E_INVALID;
    if (net && net->ctx->device != g->ctx->device)
        return arc::fail(g->ctx, ARC_E_INVALID, "cutnet on another device");
    g->net = net;
    return ARC_OK;
}

int arc_gather_set_estimate_pool(arc_gather* g, int chunks_per_tile) {
    if (!g) return ARC_E_INVALID;
    if (chunks_per_tile < 0 || chunks_per_tile > 256)
        return arc::fail(g->ctx, ARC_E_INVALID, "estimate pool: 0..256 chunks per tile");
    g->est_chunks = chunks_per_tile;
    return ARC_OK;
}

int arc_gather_set_launch_shape(arc_gather* g, int cut_stack, int anti_chunk) {
    if (!g) return ARC_E_INVALID;
    if (cut_stack != 64 && cut_stack != 96 && cut_stack != 128)
        return arc::fail(g->ctx, ARC_E_INVALID, "launch shape: cut_stack is 64, 96 or 128");
    if (anti_chunk != 768 && anti_chunk != 1536 && anti_chunk != 3072)
        return arc::fail(g->ctx, ARC_E_INVALID, "launch shape: anti_chunk is 768, 1536 or 3072");
    g->cut_stack = cut_stack;
    g->anti_chunk = anti_chunk;
    return ARC_OK;
}

int arc_gather_cut_samples(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                           const arc_gather_params* p, int flags, float* X, float* y, int capacity, uint64_t seed,
                           float keep, int* count) {
    (void)camPos;
    ARC_TRY(validate_params(g, avpls, p));
    arc_ctx* ctx = g->ctx;
    if (!X || !y || capacity < 0 || !count || !(keep > 0.f && keep <= 1.f)) return ARC_E_INVALID;
    if (g->general) return arc::fail(ctx, ARC_E_UNSUPPORTED, "the learned cut needs Lambertian materials");
    if (p->cutEstimateDepth < 0 || p->cutEstimateDepth > 40) return arc::fail(ctx, ARC_E_INVALID, "cutEstimateDepth");
    if (!(p->cutImageFloor >= 0.f)) return arc::fail(ctx, ARC_E_INVALID, "cutImageFloor");
    if (p->cutEvalDepth < 0 || p->cutEvalDepth > 4) return arc::fail(ctx, ARC_E_INVALID, "cutEvalDepth");
    *count = 0;
    const int n = avpls->numAvpls;
    if (n < 2) return ARC_OK;
    arc::AvplsDev a{};
    ARC_TRY(stage_avpls(g, avpls, flags, a));
    ARC_TRY(build_cut_tree(g, a, (flags & ARC_TRUST_INPUT) == 0));
    ARC_TRY(arc::ensure(ctx, g->counters, sizeof(unsigned long long) * C_NUM));
    int* dcount = reinterpret_cast<int*>(g->counters.p);
    ARC_HIP(ctx, hipMemsetAsync(dcount, 0, sizeof(int), ctx->stream));
    const arc::CutEBParams cp{p->errorBound, p->cutEstimateDepth, 1e-3f * sceneExtent, p->cutBudget,
                              1, p->cutCriterion, p->cutImageFloor, p->cutConfidence, 0};
    ARC_TRY(arc::cut_samples(g, g->bvh, cp, X, y, capacity, seed, keep, dcount));
    ARC_HIP(ctx, hipMemcpyAsync(count, dcount, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

int arc_gather_exact_f64(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                         const arc_gather_params* p, int flags, double* radiance3, double* antiradiance3) {
    ARC_TRY(validate_params(g, avpls, p));
    arc_ctx* ctx = g->ctx;
    const int npix = g->width * g->height;
    const int n = avpls->numAvpls;
    const float radius = p->photonRadiusScale * sceneExtent / 100.f;
    const double cos_half = std::cos((double)p->coneHalfAngle);
    arc::AvplsDev a{};
    ARC_TRY(stage_avpls(g, avpls, flags, a));
    const int splits = pick_splits(npix, std::max(n, 1));
    const int chunk = (std::max(n, 1) + splits - 1) / splits;
    ARC_TRY(arc::ensure(ctx, g->part, sizeof(double) * 6 * (size_t)splits * npix));
    if (n > 0) {
        if (p->antiradianceModel == ARC_ANTIRAD_DISK)
            hipLaunchKernelGGL(k_exact_f64<ARC_ANTIRAD_DISK>, dim3(blocks(npix), splits), dim3(kThreads), 0,
                               ctx->stream, a, g->gpos, g->gnrm, g->galb, g->mats.as<arc_material>(), npix,
                               camPos.x, camPos.y, camPos.z, radius, cos_half, chunk, g->part.as<double>());
        else
            hipLaunchKernelGGL(k_exact_f64<ARC_ANTIRAD_CONE>, dim3(blocks(npix), splits), dim3(kThreads), 0,
                               ctx->stream, a, g->gpos, g->gnrm, g->galb, g->mats.as<arc_material>(), npix,
                               camPos.x, camPos.y, camPos.z, radius, cos_half, chunk, g->part.as<double>());
        ARC_LAUNCHED(ctx);
    }
    hipLaunchKernelGGL(k_finalize_f64, dim3(blocks(npix)), dim3(kThreads), 0, ctx->stream, g->part.as<double>(),
                       n > 0 ? splits : 0, npix, radiance3, antiradiance3);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

}  // extern "C"
