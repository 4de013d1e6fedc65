// Pixels grouped in 3D for the gather's per-group work.
//
// The reference groups visible points for its clustered antiradiance by
// screen tile and depth slice (VisiblePointsBvh, bvh.cu:986-1018, built at
// 1020-1179, rebuilt on CudaGather::rebuildVisiblePointsBvh, cudaGather.h:46),
// so that a cluster's box is compact in space.  Here the pixels are sorted by
// the 30-bit Morton code of their G-buffer position in the box of all
// positions: runs of 64 consecutive sorted pixels (one wave) and of 1024 (a
// cluster) are compact in 3D on any scene, including screen tiles that
// straddle depth discontinuities.  The light cut shares one cut per run of
// 64 (the union of the pixels' refinements is small when the pixels are
// close) and the clustered antiradiance lists AVPLs per run of 1024.
//
// The order is built when the G-buffer is set and on
// arc_gather_rebuild_visible_points.  It only groups pixels: every pixel's
// result is decided by its own position and normal, so an order left stale by
// a G-buffer changed in place costs speed, never correctness.
#include <algorithm>

#include "gather_common.hpp"

namespace {

constexpr int kPT = 256;
constexpr int kPartials = 1024;

__global__ __launch_bounds__(kPT) void k_pix_bounds(const float4* __restrict__ gpos, int npix,
                                                    float* __restrict__ part) {
    __shared__ float s[6][kPT];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * kPT + threadIdx.x; i < npix; i += gridDim.x * kPT) {
        const float4 p = gpos[i];
        v[0] = fminf(v[0], p.x); v[1] = fminf(v[1], p.y); v[2] = fminf(v[2], p.z);
        v[3] = fmaxf(v[3], p.x); v[4] = fmaxf(v[4], p.y); v[5] = fmaxf(v[5], p.z);
    }
    for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kPT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + off])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Morton keys of the pixel positions in the box of the partial boxes.
__global__ __launch_bounds__(kPT) void k_pix_keys(const float4* __restrict__ gpos, int npix,
                                                  const float* __restrict__ part, int nparts,
                                                  uint32_t* __restrict__ keys, int* __restrict__ vals) {
    __shared__ float box[6];
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = k < 3 ? INFINITY : -INFINITY;
        for (int b = 0; b < nparts; ++b) v = k < 3 ? fminf(v, part[6 * b + k]) : fmaxf(v, part[6 * b + k]);
        box[k] = v;
    }
    __syncthreads();
    const int i = blockIdx.x * kPT + threadIdx.x;
    if (i >= npix) return;
    const float4 p = gpos[i];
    uint32_t q[3];
    const float c[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
        const float ext = box[3 + k] - box[k];
        const float u = ext > 0.f ? (c[k] - box[k]) / ext : 0.f;
        q[k] = (uint32_t)fminf(fmaxf(u * 1024.f, 0.f), 1023.f);  // NaN -> 0
    }
    keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
    vals[i] = i;
}

}  // namespace

namespace arc {

int pixel_order_build(arc_gather* g) {
    arc_ctx* ctx = g->ctx;
    const int npix = g->width * g->height;
    g->porder_ready = false;
    const int nparts = std::max(1, std::min(kPartials, (npix + kPT - 1) / kPT));
    ARC_TRY(ensure(ctx, g->porder, sizeof(int) * (size_t)npix));
    ARC_TRY(ensure(ctx, g->pkeys, sizeof(uint32_t) * (size_t)npix + sizeof(float) * 6 * (size_t)nparts));
    uint32_t* keys = g->pkeys.as<uint32_t>();
    float* part = reinterpret_cast<float*>(keys + npix);
    hipLaunchKernelGGL(k_pix_bounds, dim3(nparts), dim3(kPT), 0, ctx->stream, g->gpos, npix, part);
    hipLaunchKernelGGL(k_pix_keys, dim3((npix + kPT - 1) / kPT), dim3(kPT), 0, ctx->stream, g->gpos, npix, part,
                       nparts, keys, g->porder.as<int>());
    ARC_LAUNCHED(ctx);
    ARC_TRY(sort_pairs(ctx, g->psort, keys, g->porder.as<int>(), npix, 30));
    g->porder_ready = true;
    return ARC_OK;
}

}  // namespace arc
