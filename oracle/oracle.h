/*
 * oracle.h — CPU restatement of the AntiradianceCuts hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline; the product path never calls it.
 *
 * Every function names the reference lines it restates.  Two numeric modes:
 *   REF32 — FP32 with IEEE-754 division and square root, in the operation
 *           order the reference writes (compiled with -ffp-contract=off);
 *   F64   — the same formulas in double precision.
 * What pins each function is listed at its declaration; the tests in
 * tests/test_oracle_*.py carry the checks.
 */
#ifndef ARC_ORACLE_H_
#define ARC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BvhNode (bvh.h:21-28): 36 bytes, 4-byte aligned. */
typedef struct orc_node {
    int left, right;
    float bbMin[3];
    float bbMax[3];
    int visited;
} orc_node;

/* ---- morton.h ---------------------------------------------------------
 * Pinned by: exhaustive comparison of all 1024 inputs against a naive bit
 * loop, closed-form known answers, the reference header compiled in place
 * (oracle/_ref) where that builds. */
uint32_t orc_expand_bits_by1(uint32_t x);                        /* morton.h:24-32 */
uint32_t orc_expand_bits_by2(uint32_t x);                        /* morton.h:34-42 */
uint32_t orc_interleave3(uint32_t x, uint32_t y, uint32_t z);    /* morton.h:48-50 */
uint64_t orc_interleave6(uint64_t x, uint64_t y, uint64_t z,
                         uint64_t u, uint64_t v, uint64_t w);    /* morton.h:67-80 */
uint32_t orc_morton3d(float x, float y, float z);                /* morton.h:82-92 */
uint64_t orc_morton6d(float ax, float ay, float az,
                      float bx, float by, float bz);             /* morton.h:94-107 */

/* ---- bvh.cu ----------------------------------------------------------- */
/* Per-axis min / max with the reference's init values FLT_MAX and FLT_MIN
 * (bvh.cu:346-350).  Pinned by brute force and an all-negative-axis case. */
void orc_bounds(const float* p3, int n, float* mn3, float* mx3);
/* (v - min) / (max - min) per component, IEEE division (bvh.cu:73-81). */
void orc_normalize(const float* p3, int n, const float* mn3, const float* mx3, float* out3);
/* delta(i, j) of bvh.cu:124-135 over sorted codes; num_nodes = n - 1. */
int orc_delta(int i, int j, const uint32_t* codes, int num_nodes);

/* Karras radix tree over sorted codes (bvh.cu:188-266), all n - 1 internal
 * nodes; ranges[2k..2k+1] = [min(i,j), max(i,j)] of node k (not stored by
 * the reference, bvh.cu:222).  Pinned by an independent top-down builder,
 * the checkTreeIntegrity invariants and the Karras range properties. */
int orc_radix_tree(const uint32_t* codes, int n, orc_node* nodes, int* parents, int* ranges);

/* Bottom-up AABB refit (kernel_innerNodes, bvh.cu:280-316), run as the
 * sequential leaf order 0..n-1.  Pinned by brute-force min/max over each
 * node's leaf range. */
int orc_refit(orc_node* nodes, const int* parents, const float* sorted_pos3, int n);

/* checkTreeIntegrity (bvh.cu:568-639).  Returns 0 when every node is visited
 * exactly once, every internal node has two children and parent pointers
 * agree; *max_stack receives the largest DFS stack pointer. */
int orc_check_tree_integrity(const orc_node* nodes, const int* parents, int n, int* max_stack);

/* Full SimpleBvh build (Bvh::create, bvh.cu:325-454 with SimpleBvh::sort,
 * bvh.cu:886-895): bounds, normalise, Morton (3D, or 6D truncated to 32 bits
 * when consider_normals, bvh.h:17), stable sort, radix tree, refit.
 * Returns -1 when n <= 1 (bvh.cu:327-330). */
int orc_lbvh_build(const float* pos3, const float* nrm3, int n, int consider_normals,
                   uint32_t* codes, int* perm, orc_node* nodes, int* parents, int* ranges,
                   float* sorted_pos3, float* sorted_nrm3);

/* AvplBvh (bvh.cu:673-808): as orc_lbvh_build plus the node data over 2n-1
 * slots (leaves [0,n) permuted by the sort, internal node k at n + k) and the
 * cluster representatives of bvh.cu:726-744 drawn with rand[slot].  Every
 * AVPL must have bounce > 0 (returns -2 otherwise, quirk Q2).  Pinned by a
 * direct restatement of cluster() in the tests and the refit checks. */
int orc_avpl_bvh_build(int n, const float* pos3, const float* nrm3, const float* dir3,
                       const float* rad3, const int* mat, const int* bounce, const float* rand,
                       uint32_t* codes, int* perm, orc_node* nodes, int* parents, int* ranges,
                       float* sorted_pos3, float* sorted_nrm3,
                       int* nd_size, int* nd_mat, float* nd_pos3, float* nd_nrm3,
                       float* nd_rad3, float* nd_dir3);

/* ---- cudaGather.cu ------------------------------------------------------ */
enum { ORC_ANTIRAD_DISK = 0, ORC_ANTIRAD_CONE = 1 };

typedef struct orc_scene {
    int width, height;
    const float* pos4;     /* G-buffer (x, y, z, 1)                  */
    const float* nrm4;     /* G-buffer (nx, ny, nz, materialIndex)  */
    const float* albedo4;  /* optional per-pixel diffuse override   */
    const float* mats;     /* 10 floats per MATERIAL                */
    int num_mats;
    int num_avpls;
    const float *apos3, *anrm3, *adir3, *arad3, *aanti3;
    const int* amat;
    float cam[3];
    float photon_radius;   /* photonRadiusScale * sceneExtent / 100 (cudaGather.cu:949) */
    int antirad_model;     /* ORC_ANTIRAD_DISK | ORC_ANTIRAD_CONE   */
    float cone_half_angle; /* radians, CONE model                   */
} orc_scene;

/* Pair functions at one (pixel, AVPL); precision 32 (REF32) or 64 (F64).
 * getRadiance cudaGather.cu:78-95, getAntiradiance :109-144 (DISK) or
 * Shaders/Gather.frag:149-188 times the pixel BRDF (CONE).  Pinned by closed
 * forms (head-on pairs, in/out-of-disk hits, back-facing pairs). */
int orc_pair(const orc_scene* s, int precision, int pixel, int avpl, double* rad3, double* anti3);

/* Exact gather (kernel_radiance_simple :269-314 and kernel_antiradiance
 * :446-557), summed in AVPL index order, over the listed pixels (all pixels
 * when pixels == NULL).  Outputs 3 doubles per listed pixel.  Pinned by the
 * pair checks plus superposition and linearity in L and A. */
int orc_gather_exact(const orc_scene* s, int precision, const int* pixels, int npix,
                     double* rad3, double* anti3);
/* As orc_gather_exact, summing only the radiance (what & 1) and / or the
 * antiradiance (what & 2); the other output is left zero. */
int orc_gather_exact_parts(const orc_scene* s, int precision, const int* pixels, int npix,
                           double* rad3, double* anti3, int what);

/* Light-cut radiance (kernel_bvh, cudaGather.cu:316-444) over an AvplBvh:
 * stack DFS (32 entries, right child first), corner visibility cull,
 * solid-angle test against ref_thresh.  Decision arithmetic follows REF32 (or
 * F64 with precision 64).  err[p] = 1 where the reference writes its error
 * sentinel.  Counters (may be NULL) receive visited inner nodes and evaluated
 * (pixel, node) pairs.  Pinned by the ref_thresh -> 0 and ref_thresh -> 2*pi
 * limits against orc_gather_exact. */
int orc_gather_cut_solid_angle(const orc_scene* s, int precision, int n_leaves,
                               const orc_node* nodes, const int* nd_mat, const float* nd_pos3,
                               const float* nd_nrm3, const float* nd_rad3, const float* nd_dir3,
                               float ref_thresh, const int* pixels, int npix,
                               double* rad3, int* err, long long* visited, long long* evaluated);

/* ---- output side (Render.cpp, CReferenceImage.cpp) ----------------------
 * Renderer::Add (Render.cpp:788-806): accum.rgb += frame.rgb, accum.w += 1,
 * in FP32.  Renderer::Normalize (Render.cpp:642-667): out.rgb = accum.rgb *
 * (1.f / paths) (0 when paths == 0), out.w = 1.  Pinned by exact FP32
 * arithmetic (numpy float32 gives the same bits). */
void orc_image_accumulate(float* accum4, const float* frame4, int npix);
void orc_image_normalize(float* out4, const float* accum4, int npix, int num_paths);
/* CReferenceImage::GetError (CReferenceImage.cpp:28-51) as written: FP32
 * per-channel sums of min(100, d^2) in pixel order, scaled by 1.f/(w*h),
 * square roots, then (e.x * e.y + e.z) / 3.  e3 receives the per-channel
 * values.  Pinned by closed forms (constant differences, the clamp at 100,
 * the product-plus-sum combination). */
float orc_image_error(const float* ref4, const float* img4, int npix, float* e3);

/* ---- ray tracing and the AVPL shooter (Triangle.h, AvplShooter.h) -------
 * Triangle::intersect, FRONT_FACE mode (Triangle.h): e1 = p2 - p0,
 * e2 = p1 - p0, Moller-Trumbore on unnormalised barycentrics, |det| <
 * det_eps rejected (the reference's EPS = 1e-4), FP32 in glm's operation
 * order.  Returns 1 and *t on a hit (any sign of t).  Pinned by closed forms
 * (head-on hit distance, back face, edge and outside rays, the det floor). */
int orc_tri_intersect(const float* p0, const float* p1, const float* p2, const float* org, const float* dir,
                      float det_eps, float* t);
/* The closest hit with t > 0 over all triangles by brute force (ties: the
 * lower index); a triangle whose material is flagged emissive ends the ray
 * with no hit (Scene.cpp:247-252).  Returns the triangle or -1. */
int orc_trace_closest(int ntri, const float* p0, const float* p1, const float* p2, const int* material,
                      const int* emissive, const float* org, const float* dir, float det_eps, float* t);
/* AvplShooter::createPath (AvplShooter.h:77-142) for paths [first,
 * first + count) with the library's counter-based random numbers, brute-force
 * intersection.  light: center3, front3, s3, u3 (the quad's unit side axes),
 * half_w, half_h, radiance3, material (as floats).  mats: 10 floats per
 * material.  Writes each path's indirect AVPL count to counts[] and, when
 * pos3 is not NULL, its AVPLs consecutively (at most max_avpls); returns the
 * total count. */
int orc_trace_shoot(int ntri, const float* p0, const float* p1, const float* p2, const int* material,
                    const float* mats, int num_mats, const float* light, uint64_t seed, int first, int count,
                    int explicit_direct, int* counts, float* pos3, float* nrm3, float* dir3, float* rad3,
                    float* anti3, int* mat, int* bounce, int max_avpls);

/* Worker threads of the oracle's parallel loops (OpenMP) from now on; returns
 * the count in effect.  n <= 0 leaves it unchanged. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif /* ARC_ORACLE_H_ */
