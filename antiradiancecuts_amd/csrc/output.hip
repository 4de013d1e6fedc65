// Output side of the gather on gfx950: progressive accumulation, path-count
// normalisation, the reference-image error metric, and the PFM export.
//
//   arc_image_accumulate  Renderer::Add(target, source) (Render.cpp:788-806):
//                         additive blend (GL_ONE, GL_ONE) of Add.frag's output,
//                         source.rgb with w = 1, so accum.rgb += frame.rgb and
//                         accum.w += 1 (the number of frames added);
//   arc_image_normalize   Renderer::Normalize (Render.cpp:642-667) with
//                         Normalize.frag: out.rgb = accum.rgb * (1 / paths),
//                         the factor formed on the host in FP32 and 0 when
//                         paths == 0; out.w = 1;
//   arc_image_error       CReferenceImage::GetError (CReferenceImage.cpp:28-51):
//                         per channel sqrt(mean(min(100, d^2))), returned as
//                         (e.x * e.y + e.z) / 3 (the reference's formula,
//                         kept as is), plus the relative L2 and max |d| over
//                         RGB; sums in FP64 over one pass of 1024-pixel blocks;
//   arc_write_pfm /       CExport::ExportPFM (Utils/CExport.cpp:22-65): "PF",
//   arc_read_pfm          width and height, scale -1.0 (little-endian), then RGB
//                         floats in memory row order.
// One pixel per lane; all three image kernels are HBM-bound streams.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "arc_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kErrBlock = 1024;  // pixels per block of the error reduction

__global__ __launch_bounds__(kThreads) void k_accumulate(float4* __restrict__ accum, const float4* __restrict__ frame,
                                                         int npix) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= npix) return;
    const float4 a = accum[i], f = frame[i];
    accum[i] = make_float4(a.x + f.x, a.y + f.y, a.z + f.z, a.w + 1.f);
}

__global__ __launch_bounds__(kThreads) void k_normalize(float4* __restrict__ out, const float4* __restrict__ accum,
                                                        int npix, float inv) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= npix) return;
    const float4 a = accum[i];
    out[i] = make_float4(a.x * inv, a.y * inv, a.z * inv, 1.f);
}

// Per block: sums of min(100, d^2) per channel, of d^2 and ref^2 over RGB,
// and max |d|; 6 doubles per block.
__global__ __launch_bounds__(kThreads) void k_error_partial(const float4* __restrict__ ref,
                                                            const float4* __restrict__ img, int npix,
                                                            double* __restrict__ partial) {
    __shared__ double sh[6][kThreads];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int base = blockIdx.x * kErrBlock;
    for (int k = threadIdx.x; k < kErrBlock; k += kThreads) {
        const int i = base + k;
        if (i >= npix) break;
        const float4 r = ref[i], u = img[i];
        const float dx = r.x - u.x, dy = r.y - u.y, dz = r.z - u.z;
        acc[0] += (double)fminf(100.f, dx * dx);
        acc[1] += (double)fminf(100.f, dy * dy);
        acc[2] += (double)fminf(100.f, dz * dz);
        // the reference's clamped terms use the FP32 differences; the
        // relative L2 and max |d| use exact (FP64) ones
        const double ex = (double)r.x - u.x, ey = (double)r.y - u.y, ez = (double)r.z - u.z;
        acc[3] += ex * ex + ey * ey + ez * ez;
        acc[4] += (double)r.x * r.x + (double)r.y * r.y + (double)r.z * r.z;
        acc[5] = fmax(acc[5], fmax(fabs(ex), fmax(fabs(ey), fabs(ez))));
    }
    for (int c = 0; c < 6; ++c) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 6; ++c)
                sh[c][threadIdx.x] = c == 5 ? fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + off])
                                            : sh[c][threadIdx.x] + sh[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// One block folds the partials in block order (deterministic).
__global__ __launch_bounds__(64) void k_error_final(const double* __restrict__ partial, int nblocks,
                                                    double* __restrict__ out) {
    if (threadIdx.x >= 6) return;
    const int c = threadIdx.x;
    double a = 0.0;
    for (int b = 0; b < nblocks; ++b) a = c == 5 ? fmax(a, partial[6 * (size_t)b + c]) : a + partial[6 * (size_t)b + c];
    out[c] = a;
}

int blocks(int n, int per) { return (n + per - 1) / per; }

}  // namespace

extern "C" {

int arc_image_accumulate(arc_ctx* ctx, float* accum4, const float* frame4, int npix) {
    if (!ctx || npix < 0 || (npix > 0 && (!accum4 || !frame4))) return ARC_E_INVALID;
    if (npix == 0) return ARC_OK;
    hipLaunchKernelGGL(k_accumulate, dim3(blocks(npix, kThreads)), dim3(kThreads), 0, ctx->stream,
                       reinterpret_cast<float4*>(accum4), reinterpret_cast<const float4*>(frame4), npix);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int arc_image_normalize(arc_ctx* ctx, float* out4, const float* accum4, int npix, int num_paths) {
    if (!ctx || npix < 0 || num_paths < 0 || (npix > 0 && (!out4 || !accum4))) return ARC_E_INVALID;
    if (npix == 0) return ARC_OK;
    const float inv = num_paths != 0 ? 1.f / (float)num_paths : 0.f;  // Render.cpp:651-654
    hipLaunchKernelGGL(k_normalize, dim3(blocks(npix, kThreads)), dim3(kThreads), 0, ctx->stream,
                       reinterpret_cast<float4*>(out4), reinterpret_cast<const float4*>(accum4), npix, inv);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int arc_image_error(arc_ctx* ctx, const float* ref4, const float* img4, int npix, arc_image_error_t* out) {
    if (!ctx || !out || npix <= 0 || !ref4 || !img4) return ARC_E_INVALID;
    const int nb = blocks(npix, kErrBlock);
    double* d = nullptr;
    ARC_HIP(ctx, hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(double) * (6 * (size_t)nb + 6), ctx->stream));
    hipLaunchKernelGGL(k_error_partial, dim3(nb), dim3(kThreads), 0, ctx->stream,
                       reinterpret_cast<const float4*>(ref4), reinterpret_cast<const float4*>(img4), npix, d);
    hipLaunchKernelGGL(k_error_final, dim3(1), dim3(64), 0, ctx->stream, d, nb, d + 6 * (size_t)nb);
    double h[6];
    const hipError_t le = hipGetLastError();
    hipError_t e = le == hipSuccess ? hipMemcpyAsync(h, d + 6 * (size_t)nb, sizeof(h), hipMemcpyDeviceToHost,
                                                     ctx->stream)
                                    : le;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFreeAsync(d, ctx->stream);
    if (e != hipSuccess) return arc::hip_fail(ctx, e, "arc_image_error", __FILE__, __LINE__);
    // CReferenceImage.cpp:46-50 in FP32 from the FP64 sums
    const float inv = 1.f / (float)npix;
    const float ex = sqrtf(inv * (float)h[0]), ey = sqrtf(inv * (float)h[1]), ez = sqrtf(inv * (float)h[2]);
    out->referenceError = 1.f / 3.f * (ex * ey + ez);
    out->rmse[0] = ex;
    out->rmse[1] = ey;
    out->rmse[2] = ez;
    out->relL2 = h[4] > 0.0 ? sqrt(h[3] / h[4]) : sqrt(h[3]);
    out->maxAbs = h[5];
    return ARC_OK;
}

int arc_write_pfm(const char* path, const float* rgba_host, int width, int height) {
    if (!path || !rgba_host || width <= 0 || height <= 0) return ARC_E_INVALID;
    FILE* f = std::fopen(path, "wb");
    if (!f) return ARC_E_INVALID;
    const std::string head = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    std::vector<float> row(3 * (size_t)width);
    for (int y = 0; ok && y < height; ++y) {
        for (int x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) row[3 * (size_t)x + c] = rgba_host[4 * ((size_t)y * width + x) + c];
        ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? ARC_OK : ARC_E_INVALID;
}

int arc_read_pfm(const char* path, float* rgba_host, int max_pixels, int* width, int* height) {
    if (!path || !width || !height) return ARC_E_INVALID;
    FILE* f = std::fopen(path, "rb");
    if (!f) return ARC_E_INVALID;
    char tag[3] = {0, 0, 0};
    int w = 0, h = 0;
    float scale = 0.f;
    const bool head = std::fread(tag, 1, 2, f) == 2 && tag[0] == 'P' && tag[1] == 'F' &&
                      std::fscanf(f, "%d %d %f", &w, &h, &scale) == 3 && std::fgetc(f) == '\n' && w > 0 && h > 0 &&
                      scale < 0.f;  // little-endian only
    if (!head) {
        std::fclose(f);
        return ARC_E_INVALID;
    }
    *width = w;
    *height = h;
    if (!rgba_host) {  // size query
        std::fclose(f);
        return ARC_OK;
    }
    if ((long long)w * h > max_pixels) {
        std::fclose(f);
        return ARC_E_INVALID;
    }
    std::vector<float> row(3 * (size_t)w);
    bool ok = true;
    for (int y = 0; ok && y < h; ++y) {
        ok = std::fread(row.data(), sizeof(float), row.size(), f) == row.size();
        for (int x = 0; ok && x < w; ++x) {
            float* o = rgba_host + 4 * ((size_t)y * w + x);
            o[0] = row[3 * (size_t)x];
            o[1] = row[3 * (size_t)x + 1];
            o[2] = row[3 * (size_t)x + 2];
            o[3] = 1.f;
        }
    }
    std::fclose(f);
    return ok ? ARC_OK : ARC_E_INVALID;
}

}  // extern "C"
