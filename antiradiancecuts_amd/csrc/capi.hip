// C ABI glue: contexts, status strings, device memory helpers and the Bvh
// entry points (bvh.h:53-161 surface).
#include <cstdio>
#include <cstring>
#include <vector>

#include "arc_internal.hpp"

namespace arc {

int fail(arc_ctx* ctx, int status, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return status;
}

int hip_fail(arc_ctx* ctx, hipError_t e, const char* expr, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), file, line);
    if (ctx) ctx->last_error = buf;
    return e == hipErrorOutOfMemory ? ARC_E_NOMEM : ARC_E_HIP;
}

int ensure(arc_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return ARC_OK;
    if (bytes == 0) bytes = 4;
    // grow by 25% headroom so per-frame size jitter does not reallocate
    const size_t want = b.bytes ? std::max(bytes, b.bytes + b.bytes / 4) : bytes;
    if (b.p) {
        // the stream may still use the old block
        ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ARC_HIP(ctx, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    ARC_HIP(ctx, hipMalloc(&b.p, want));
    b.bytes = want;
    return ARC_OK;
}

void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

}  // namespace arc

using arc::f3;

extern "C" {

const char* arc_version(void) { return "arc 0.1.0 gfx950"; }

const char* arc_status_string(int s) {
    switch (s) {
        case ARC_OK: return "ok";
        case ARC_E_INVALID: return "invalid argument";
        case ARC_E_TOO_FEW_POINTS: return "not enough points for bvh construction";
        case ARC_E_HIP: return "HIP runtime error";
        case ARC_E_NOMEM: return "out of device memory";
        case ARC_E_NO_DEVICE: return "no HIP device";
        case ARC_E_UNSUPPORTED: return "unsupported combination";
        case ARC_E_STATE: return "object not ready";
        default: return "unknown status";
    }
}

int arc_ctx_create(int device, arc_ctx** out) {
    if (!out) return ARC_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ARC_E_NO_DEVICE;
    if (device < 0 || device >= count) return ARC_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return ARC_E_HIP;
    arc_ctx* c = new arc_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ARC_E_HIP;
    }
    c->stream = c->own_stream;
    *out = c;
    return ARC_OK;
}

void arc_ctx_destroy(arc_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) {
        (void)hipStreamSynchronize(ctx->own_stream);
        (void)hipStreamDestroy(ctx->own_stream);
    }
    delete ctx;
}

int arc_ctx_set_stream(arc_ctx* ctx, void* s) {
    if (!ctx) return ARC_E_INVALID;
    ctx->stream = static_cast<hipStream_t>(s);  // NULL is the device's default (null) stream
    return ARC_OK;
}

int arc_ctx_use_own_stream(arc_ctx* ctx) {
    if (!ctx) return ARC_E_INVALID;
    ctx->stream = ctx->own_stream;
    return ARC_OK;
}

void* arc_ctx_get_stream(arc_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int arc_ctx_synchronize(arc_ctx* ctx) {
    if (!ctx) return ARC_E_INVALID;
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

const char* arc_ctx_last_error(arc_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int arc_device_alloc(arc_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return ARC_E_INVALID;
    ARC_HIP(ctx, hipMalloc(out, bytes ? bytes : 4));
    return ARC_OK;
}

int arc_device_free(arc_ctx* ctx, void* p) {
    if (!ctx) return ARC_E_INVALID;
    if (p) ARC_HIP(ctx, hipFree(p));
    return ARC_OK;
}

int arc_copy_to_device(arc_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return ARC_E_INVALID;
    ARC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

int arc_copy_to_host(arc_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return ARC_E_INVALID;
    ARC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

// ------------------------------------------------------------------ Bvh --
int arc_bvh_create(arc_ctx* ctx, int consider_normals, arc_bvh** out) {
    if (!ctx || !out) return ARC_E_INVALID;
    arc_bvh* b = new arc_bvh();
    b->ctx = ctx;
    b->consider_normals = consider_normals ? 1 : 0;
    *out = b;
    return ARC_OK;
}

void arc_bvh_destroy(arc_bvh* b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    arc::DevBuf* bufs[] = {&b->codes,   &b->perm,    &b->nodes,  &b->parents, &b->ranges,   &b->positions,
                           &b->normals, &b->in_pos,  &b->in_nrm, &b->nd_size, &b->nd_mat,   &b->nd_rand,
                           &b->nd_pos,  &b->nd_nrm,  &b->nd_rad, &b->nd_dir,  &b->nd_merge, &b->staged, &b->bounds,
                           &b->partials, &b->ready, &b->boxes, &b->sort.keys_alt, &b->sort.vals_alt,
                           &b->sort.counts};
    for (arc::DevBuf* d : bufs) arc::release(*d);
    delete b;
}

int arc_bvh_build_points(arc_bvh* b, const arc_float3* positions, const arc_float3* normals, int n, int flags) {
    if (!b || (n > 0 && !positions)) return ARC_E_INVALID;
    arc_ctx* ctx = b->ctx;
    const f3* pos = reinterpret_cast<const f3*>(positions);
    const f3* nrm = reinterpret_cast<const f3*>(normals);
    if ((flags & ARC_INPUT_HOST) && n > 0) {
        const size_t bytes = sizeof(f3) * (size_t)n;
        ARC_TRY(arc::ensure(ctx, b->in_pos, bytes));
        ARC_HIP(ctx, hipMemcpyAsync(b->in_pos.p, positions, bytes, hipMemcpyHostToDevice, ctx->stream));
        pos = b->in_pos.as<f3>();
        if (normals) {
            ARC_TRY(arc::ensure(ctx, b->in_nrm, bytes));
            ARC_HIP(ctx, hipMemcpyAsync(b->in_nrm.p, normals, bytes, hipMemcpyHostToDevice, ctx->stream));
            nrm = b->in_nrm.as<f3>();
        }
        ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // host arrays are free to reuse on return
    }
    return arc::bvh_build_points(b, pos, nrm, n);
}

int arc_bvh_build_avpls(arc_bvh* b, const arc_avpls* a, const float* rand, uint64_t seed, int flags) {
    if (!b || !a || a->numAvpls < 0) return ARC_E_INVALID;
    arc_ctx* ctx = b->ctx;
    const int n = a->numAvpls;
    if (n > 0 && (!a->position || !a->normal || !a->incDirection || !a->incRadiance || !a->materialIndex ||
                  !a->bounce))
        return arc::fail(ctx, ARC_E_INVALID, "AVPL arrays missing");
    arc::AvplsDev d{reinterpret_cast<const f3*>(a->position), reinterpret_cast<const f3*>(a->normal),
                    reinterpret_cast<const f3*>(a->incDirection), reinterpret_cast<const f3*>(a->incRadiance),
                    reinterpret_cast<const f3*>(a->antiradiance), a->materialIndex, a->bounce, n};
    const float* r = rand;
    if ((flags & ARC_INPUT_HOST) && n > 0) {
        // stage the SoA (and rand) in the in_pos buffer: 5 float3 + 2 int arrays + 2n-1 floats
        const size_t v3 = sizeof(f3) * (size_t)n, vi = sizeof(int) * (size_t)n;
        const size_t vr = rand ? sizeof(float) * (size_t)(2 * n - 1) : 0;
        ARC_TRY(arc::ensure(ctx, b->in_pos, 5 * v3 + 2 * vi + vr));
        char* base = b->in_pos.as<char>();
        const void* src[8] = {a->position, a->normal, a->incDirection, a->incRadiance, a->antiradiance,
                              a->materialIndex, a->bounce, rand};
        const size_t sz[8] = {v3, v3, v3, v3, v3, vi, vi, vr};
        char* dst[8];
        size_t off = 0;
        for (int k = 0; k < 8; ++k) {
            dst[k] = base + off;
            off += sz[k];
            if (src[k] && sz[k])
                ARC_HIP(ctx, hipMemcpyAsync(dst[k], src[k], sz[k], hipMemcpyHostToDevice, ctx->stream));
        }
        ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // host arrays are free to reuse on return
        d.pos = reinterpret_cast<const f3*>(dst[0]);
        d.nrm = reinterpret_cast<const f3*>(dst[1]);
        d.dir = reinterpret_cast<const f3*>(dst[2]);
        d.rad = reinterpret_cast<const f3*>(dst[3]);
        d.anti = a->antiradiance ? reinterpret_cast<const f3*>(dst[4]) : nullptr;
        d.mat = reinterpret_cast<const int*>(dst[5]);
        d.bounce = reinterpret_cast<const int*>(dst[6]);
        r = rand ? reinterpret_cast<const float*>(dst[7]) : nullptr;
    }
    return arc::bvh_build_avpls(b, d, r, seed, (flags & ARC_TRUST_INPUT) == 0);
}

int arc_bvh_random_numbers(arc_ctx* ctx, uint64_t seed, int count, float* out) {
    if (!ctx || count < 0 || (count > 0 && !out)) return ARC_E_INVALID;
    return arc::random_numbers(ctx, seed, count, out);
}

int arc_bvh_get_param(arc_bvh* b, arc_bvh_param* out) {
    if (!b || !out) return ARC_E_INVALID;
    if (b->n <= 1) return arc::fail(b->ctx, ARC_E_STATE, "bvh not built");
    out->numLeafs = b->n;
    out->numNodes = b->n - 1;
    out->nodes = b->nodes.as<arc_bvh_node>();
    out->positions = b->positions.as<arc_float3>();
    out->normals = b->normals.as<arc_float3>();
    return ARC_OK;
}

int arc_bvh_get_data(arc_bvh* b, arc_bvh_data* out) {
    if (!b || !out) return ARC_E_INVALID;
    if (b->n <= 1) return arc::fail(b->ctx, ARC_E_STATE, "bvh not built");
    out->morton = b->codes.as<uint32_t>();
    out->parents = b->parents.as<int>();
    out->numLeafs = b->n;
    out->numNodes = b->n - 1;
    out->perm = b->perm.as<int>();
    out->ranges = b->ranges.as<int>();
    return ARC_OK;
}

int arc_bvh_get_node_data(arc_bvh* b, arc_avpl_node_data* out) {
    if (!b || !out) return ARC_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    if (b->n <= 1) return arc::fail(b->ctx, ARC_E_STATE, "bvh not built");
    if (!b->has_node_data) return ARC_OK;
    out->size = b->nd_size.as<int>();
    out->materialIndex = b->nd_mat.as<int>();
    out->randomNumbers = b->nd_rand.as<float>();
    out->position = b->nd_pos.as<arc_float3>();
    out->normal = b->nd_nrm.as<arc_float3>();
    out->incRadiance = b->nd_rad.as<arc_float3>();
    out->incDirection = b->nd_dir.as<arc_float3>();
    return ARC_OK;
}

int arc_bvh_get_node(arc_bvh* b, int i, arc_bvh_node* out) {
    if (!b || !out) return ARC_E_INVALID;
    if (b->n <= 1) return arc::fail(b->ctx, ARC_E_STATE, "bvh not built");
    if (i < 0 || i >= b->n - 1) return arc::fail(b->ctx, ARC_E_INVALID, "index out of range");
    arc_ctx* ctx = b->ctx;
    ARC_HIP(ctx, hipMemcpyAsync(out, b->nodes.as<arc_bvh_node>() + i, sizeof(arc_bvh_node), hipMemcpyDeviceToHost,
                                ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

int arc_bvh_check_integrity(arc_bvh* b, int* max_stack) {
    if (!b) return ARC_E_INVALID;
    if (b->n <= 1) return arc::fail(b->ctx, ARC_E_STATE, "bvh not built");
    arc_ctx* ctx = b->ctx;
    const int n = b->n, nn = n - 1;
    std::vector<arc_bvh_node> nodes(nn);
    std::vector<int> parents(2 * n - 1);
    ARC_HIP(ctx, hipMemcpyAsync(nodes.data(), b->nodes.p, sizeof(arc_bvh_node) * nn, hipMemcpyDeviceToHost,
                                ctx->stream));
    ARC_HIP(ctx, hipMemcpyAsync(parents.data(), b->parents.p, sizeof(int) * (2 * n - 1), hipMemcpyDeviceToHost,
                                ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // DFS from the root (internal node 0 = -1), as checkTreeIntegrity (bvh.cu:568-639)
    std::vector<int> stack{-1}, seen(2 * n - 1, 0), kids(nn, 0);
    size_t maxsp = 1;
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        int slot;
        if (node >= 0) {
            if (node >= n) return arc::fail(ctx, ARC_E_STATE, "leaf index out of range");
            slot = node;
        } else {
            const int k = -(node + 1);
            if (k >= nn) return arc::fail(ctx, ARC_E_STATE, "node index out of range");
            slot = n + k;
            stack.push_back(nodes[k].left);
            stack.push_back(nodes[k].right);
            maxsp = std::max(maxsp, stack.size());
        }
        const int par = parents[slot];
        if (par != -1) {
            if (par < 0 || par >= nn) return arc::fail(ctx, ARC_E_STATE, "parent out of range");
            kids[par]++;
        }
        if (seen[slot]++) return arc::fail(ctx, ARC_E_STATE, "node visited twice");
        if (stack.size() > (size_t)(4 * n)) return arc::fail(ctx, ARC_E_STATE, "cycle");
    }
    for (int k = 0; k < nn; ++k)
        if (kids[k] != 2) return arc::fail(ctx, ARC_E_STATE, "internal node without two children");
    for (int s = 0; s < 2 * n - 1; ++s)
        if (seen[s] != 1) return arc::fail(ctx, ARC_E_STATE, "node not reached");
    if (parents[n] != -1) return arc::fail(ctx, ARC_E_STATE, "root has a parent");
    if (max_stack) *max_stack = (int)maxsp;
    return ARC_OK;
}

}  // extern "C"
