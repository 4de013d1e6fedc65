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
    bool has_moments = false;
    arc::DevBuf codes, perm, nodes, parents, ranges, positions, normals;
    arc::DevBuf in_pos, in_nrm;                 // staged host input
    arc::DevBuf nd_size, nd_mat, nd_rand, nd_pos, nd_nrm, nd_rad, nd_dir;
    arc::DevBuf nd_radb;                        // per-slot sums of L x diffuse / pi (error-bound cut)
    arc::DevBuf nd_mom, nd_nmom;                // per-slot weighted position / normal moments
    arc::DevBuf staged;                         // AVPLs as 64-byte records for the payload permute
    arc::DevBuf bounds;                         // 12 floats + validation flag
    arc::DevBuf partials;                       // reduction partials
    arc::DevBuf ready;                          // refit: nodes spanning blocks, ready to merge
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
// moments: also sum the BRDF-weighted intensity and its position / normal
// moments per node (the error-bound cut's node estimates), using mats.
int bvh_build_avpls(arc_bvh* b, const AvplsDev& a, const float* rand, uint64_t seed, bool validate,
                    const arc_material* mats = nullptr, bool moments = false);
int random_numbers(arc_ctx* ctx, uint64_t seed, int count, float* out);
}  // namespace arc
