// Internal types shared by the HIP translation units of libarc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "arc.h"

struct arc_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
};

namespace arc {

// ------------------------------------------------------------ errors ----
int fail(arc_ctx* ctx, int status, const std::string& msg);
int hip_fail(arc_ctx* ctx, hipError_t e, const char* expr, const char* file, int line);

#define ARC_HIP(ctx, expr)                                                  \
    do {                                                                    \
        hipError_t arc_e_ = (expr);                                         \
        if (arc_e_ != hipSuccess)                                           \
            return ::arc::hip_fail((ctx), arc_e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define ARC_TRY(expr)               \
    do {                            \
        int arc_s_ = (expr);        \
        if (arc_s_ != ARC_OK)       \
            return arc_s_;          \
    } while (0)

// Launch check: picks up launch-configuration errors without synchronising.
#define ARC_LAUNCHED(ctx) ARC_HIP(ctx, hipGetLastError())

// --------------------------------------------------- device buffers ----
// A device allocation owned by a library object; grows, never shrinks.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};
int ensure(arc_ctx* ctx, DevBuf& b, size_t bytes);
void release(DevBuf& b);

// ---------------------------------------------------------- layouts ----
struct f3 {
    float x, y, z;
};
static_assert(sizeof(f3) == 12, "float3 layout");
static_assert(sizeof(arc_bvh_node) == 36, "BvhNode layout");
static_assert(sizeof(arc_material) == 40, "MATERIAL layout");

// ------------------------------------------------------------- sort ----
// Stable LSD radix sort of (uint32 key, int32 value) pairs over the low
// `key_bits` bits; the result ends in keys / vals.
struct SortScratch {
    DevBuf keys_alt, vals_alt, counts;
};
constexpr int kSortTile = 2048;   // keys per sort block (one tile)
constexpr int kSortRadix = 256;   // 8-bit digits
// the digit-major count table: kSortRadix x tiles, then the kSortRadix digit totals
constexpr size_t sort_counts_size(size_t n) {
    return (size_t)kSortRadix * ((n + kSortTile - 1) / kSortTile) + kSortRadix;
}
int sort_pairs(arc_ctx* ctx, SortScratch& s, uint32_t* keys, int* vals, int n, int key_bits);

// Counter-based uniform floats in [0,1): (mix64(seed, i) >> 40) * 2^-24.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline float hash_uniform(uint64_t seed, uint64_t i) {
    const uint64_t h = mix64(seed * 0x9e3779b97f4a7c15ULL + mix64(i + 0x632be59bd9b4e019ULL));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

}  // namespace arc

// ------------------------------------------------------------- LBVH ----
struct arc_bvh {
    arc_ctx* ctx = nullptr;
    int consider_normals = 0;
    int n = 0;               // leaves of the current build (0: none)
    bool has_node_data = false;
    bool has_cut_records = false;  // the last build was bvh_build_cut
    int refit_top_threads = 1024;  // width of the top refit (64..1024, a multiple of 64)
    arc::DevBuf codes, perm, nodes, parents, ranges, positions, normals;
    arc::DevBuf in_pos, in_nrm;                 // staged host input
    arc::DevBuf nd_size, nd_mat, nd_rand, nd_pos, nd_nrm, nd_rad, nd_dir;
    arc::DevBuf nd_merge;                       // bvh_build_cut: kMrec float4 of sums per slot
    arc::DevBuf staged;                         // AVPLs as 64-byte records for the payload permute
    arc::DevBuf bounds;                         // 12 floats + validation flag
    arc::DevBuf partials;                       // reduction partials
    arc::DevBuf ready;                          // refit: nodes spanning blocks, ready to merge
    arc::DevBuf boxes;                          // points refit: chunk and block boxes
    arc::SortScratch sort;
};

namespace arc {
// AVPL inputs on the device (possibly staged from host memory).
struct AvplsDev {
    const f3 *pos, *nrm, *dir, *rad, *anti;
    const int *mat, *bounce;
    int n;
};
int bvh_build_points(arc_bvh* b, const f3* pos, const f3* nrm, int n);
// Per internal node of the last build, the box of its leaves' unit normals
// (the light cuts' leaf records, crec slots [0, n), record 1): nbox holds
// (lo, 0), (hi, 0) per node.
int node_normal_boxes(arc_bvh* b, const float4* crec, float4* nbox);
int bvh_build_avpls(arc_bvh* b, const AvplsDev& a, const float* rand, uint64_t seed, bool validate);

// The light cuts' node records, kCrec float4 (one 64-byte line) per slot,
// leaves then internal nodes:
//   r0 = (centroid, max(I)), r1 = (m, (1 - |m|) / 4), r2 = (I, 1.2 (1 - |m|) / sqrt(count)),
//   r3 = (s^2, s, count, -)
// with I the sum of L x diffuse / pi over the node's AVPLs (the BRDF-weighted
// intensity), w = (I.x + I.y + I.z) / 3 per AVPL, the centroid sum(w x) /
// sum(w), m = sum(w n) / sum(w) over unit normals n, and s the half-diagonal
// of the node's box (0 for a leaf).  A node with sum(w) = 0 contributes
// nothing; its centroid is its AVPL's position (leaf) or its box centre and
// its m the leaf's unit normal or 0.  Everything of the cut's tests that does
// not depend on the pixel is here.
constexpr int kCrec = 4;
// Sums merged up the tree: (I, count), (sum w x, sum w), (sum w n, -).
constexpr int kMrec = 3;

__device__ inline void write_cut_record(float4* __restrict__ o, f3 I, float4 mo, f3 nm, int count, float s2, f3 fc,
                                        f3 fn) {
    f3 c, m;
    if (mo.w > 0.f) {
        const float iw = 1.0f / mo.w;
        c = f3{mo.x * iw, mo.y * iw, mo.z * iw};
        m = f3{nm.x * iw, nm.y * iw, nm.z * iw};
    } else {
        c = fc;
        const float l = fn.x * fn.x + fn.y * fn.y + fn.z * fn.z;
        const float iq = l > 0.f ? 1.0f / sqrtf(l) : 0.f;
        m = f3{fn.x * iq, fn.y * iq, fn.z * iq};
    }
    const float ml = fminf(1.f, sqrtf(m.x * m.x + m.y * m.y + m.z * m.z));
    const float spread = 1.2f * (1.f - ml) / sqrtf(fmaxf((float)count, 1.f));
    o[0] = make_float4(c.x, c.y, c.z, fmaxf(I.x, fmaxf(I.y, I.z)));
    o[1] = make_float4(m.x, m.y, m.z, 0.25f * (1.f - ml));
    o[2] = make_float4(I.x, I.y, I.z, spread);
    o[3] = make_float4(s2, sqrtf(s2), (float)count, 0.f);
}

// The LBVH of the light cuts: the same codes, sort and radix tree as
// bvh_build_avpls, but the refit sums only what the cuts read (merge records
// in nd_merge) and writes the node records to crec (kCrec x (2n - 1) float4),
// not the reference's representative node data.  mats: the 256-entry
// material table (Lambertian diffuse).
int bvh_build_cut(arc_bvh* b, const AvplsDev& a, bool validate, const arc_material* mats, float4* crec);
int random_numbers(arc_ctx* ctx, uint64_t seed, int count, float* out);
}  // namespace arc
