// The SimpleBvh points build without any device allocation.
//
// arc_bvh_build_points (Bvh::create for SimpleBvh, bvh.cu:325-454) keeps its
// outputs and its scratch (the radix sort's second key/value buffers and
// digit counts, the bounds partials, the refit's node list and its chunk and
// block boxes) in buffers it allocates and grows.  This variant writes into buffers the caller owns and
// keeps its scratch in `scratch`, a fixed device array of this file: it calls
// no hipMalloc, hipFree or allocating library routine, so callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) can build the tree.  The kernels are the ones of the
// allocating build, run over views of those buffers.
//
// Each GPU holds its own copy of `scratch` until the process ends; the
// array's device address is looked up once per device, on the first call
// there, and kept.  Every call on a GPU shares that GPU's scratch: calls that
// overlap on one GPU race.  Nothing in it has to start at zero: every kernel
// writes what it reads (the sort's counts and totals, the bounds, the node
// list and its counter, which the tree kernel clears, the chunk and block
// boxes) before reading it.
#include <mutex>

#include "arc_internal.hpp"
#include "arc_noalloc.h"

namespace {

constexpr int kMaxPoints = ARC_NOALLOC_MAX_POINTS;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
// The sort's second key/value buffers and digit count table (sort.hip);
// bounds partials of up to 1024 blocks x 12 floats (lbvh.hip).
constexpr size_t kKeysAlt = align_up(sizeof(uint32_t) * kMaxPoints);
constexpr size_t kValsAlt = align_up(sizeof(int) * kMaxPoints);
constexpr size_t kCounts = align_up(sizeof(uint32_t) * arc::sort_counts_size(kMaxPoints));
constexpr size_t kReady = align_up(sizeof(int) * 2 * (size_t)kMaxPoints);
constexpr size_t kBounds = align_up(sizeof(float) * 16);
constexpr size_t kPartials = align_up(sizeof(float) * 12 * 1024);
// the points refit's chunk (32 leaves) and block (1024 leaves) boxes, 24 bytes each
constexpr size_t kBoxes = align_up(24 * ((size_t)kMaxPoints / 32 + (size_t)kMaxPoints / 1024 + 2));
constexpr size_t kScratchBytes = kKeysAlt + kValsAlt + kCounts + kReady + kBounds + kPartials + kBoxes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of this device's `scratch` (hipGetSymbolAddress once per
// device, with the device current).
int scratch_address(arc_ctx* ctx, unsigned char** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return arc::fail(ctx, ARC_E_UNSUPPORTED, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[ctx->device]) {
        int cur = -1;
        ARC_HIP(ctx, hipGetDevice(&cur));
        if (cur != ctx->device) ARC_HIP(ctx, hipSetDevice(ctx->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != ctx->device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return arc::hip_fail(ctx, e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) return arc::fail(ctx, ARC_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch[ctx->device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch[ctx->device]);
    return ARC_OK;
}

arc::DevBuf view(void* p, size_t bytes) {
    arc::DevBuf b;
    b.p = p;
    b.bytes = bytes;
    return b;
}

}  // namespace

int arc_bvh_build_points_noalloc(arc_ctx* ctx, const arc_float3* positions, const arc_float3* normals, int n,
                                 int consider_normals, const arc_lbvh_outputs* out) {
    if (!ctx || !out || (n > 0 && !positions)) return ARC_E_INVALID;
    if (n <= 1) return arc::fail(ctx, ARC_E_TOO_FEW_POINTS, "not enough points for bvh construction");
    if (n > kMaxPoints)
        return arc::fail(ctx, ARC_E_INVALID, "arc_bvh_build_points_noalloc: more than ARC_NOALLOC_MAX_POINTS points");
    if (!out->codes || !out->perm || !out->nodes || !out->parents || !out->ranges || !out->positions)
        return arc::fail(ctx, ARC_E_INVALID, "arc_bvh_build_points_noalloc: an output buffer is missing");
    if (consider_normals && !normals) return arc::fail(ctx, ARC_E_INVALID, "normals required");
    if (normals && !out->normals)
        return arc::fail(ctx, ARC_E_INVALID, "arc_bvh_build_points_noalloc: normals given but no normals output");
    unsigned char* s = nullptr;
    ARC_TRY(scratch_address(ctx, &s));
    const size_t un = (size_t)n;
    // a build over views: every buffer is large enough, so no ensure() allocates
    arc_bvh b;
    b.ctx = ctx;
    b.consider_normals = consider_normals ? 1 : 0;
    b.codes = view(out->codes, sizeof(uint32_t) * un);
    b.perm = view(out->perm, sizeof(int) * un);
    b.nodes = view(out->nodes, sizeof(arc_bvh_node) * (un - 1));
    b.parents = view(out->parents, sizeof(int) * (2 * un - 1));
    b.ranges = view(out->ranges, sizeof(int) * 2 * (un - 1));
    b.positions = view(out->positions, sizeof(arc::f3) * un);
    // without normals the permute writes none: the positions stand in for the view's size
    b.normals = normals ? view(out->normals, sizeof(arc::f3) * un) : view(out->positions, sizeof(arc::f3) * un);
    size_t off = 0;
    b.sort.keys_alt = view(s + off, kKeysAlt);
    off += kKeysAlt;
    b.sort.vals_alt = view(s + off, kValsAlt);
    off += kValsAlt;
    b.sort.counts = view(s + off, kCounts);
    off += kCounts;
    b.ready = view(s + off, kReady);
    off += kReady;
    b.bounds = view(s + off, kBounds);
    off += kBounds;
    b.partials = view(s + off, kPartials);
    off += kPartials;
    b.boxes = view(s + off, kBoxes);
    const int st = arc::bvh_build_points(&b, reinterpret_cast<const arc::f3*>(positions),
                                         reinterpret_cast<const arc::f3*>(normals), n);
    // the views belong to the caller and to `scratch`: never released here
    return st;
}
