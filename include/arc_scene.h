/*
 * arc_scene.h — seeded synthetic inputs for the AVPL gather and the LBVH.
 *
 * The reference produces its inputs with an OpenGL G-buffer raster
 * (Render.cpp:740-750, Shaders/CreateGBuffer.frag:14-15) and a CPU path
 * tracer that shoots AVPLs (AvplShooter.h:77-142).  Neither exists on a
 * gfx950 host, so this library ray-casts a synthetic Cornell-style scene
 * into the same G-buffer layout and samples AVPLs on its surfaces:
 *
 *   G-buffer position  float4 (x, y, z, 1)
 *   G-buffer normal    float4 (nx, ny, nz, materialIndex)
 *   per-pixel albedo   float4 (r, g, b, 0)   (optional Lambertian override)
 *   AVPL SoA           float3 position / normal / incDirection /
 *                      incRadiance / antiradiance, int materialIndex / bounce
 *                      (the AvplsGpuParam layout, data_gpu.h:14-24)
 *
 * Everything is host code, deterministic for a given seed and independent of
 * the thread count (every pixel / AVPL / point draws from its own counter-
 * based stream).  All buffers are allocated by the caller.
 */
#ifndef ARC_SCENE_H_
#define ARC_SCENE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene description.  Unused fields may be zero. */
typedef struct arc_scene_desc {
    uint64_t seed;
    int width, height;          /* G-buffer resolution                        */
    int num_avpls;              /* AVPLs to sample on the surfaces            */
    int min_bounce, max_bounce; /* bounce drawn uniformly in [min, max]       */
    float radiance_lo, radiance_hi; /* per-channel incident radiance range    */
    int radiance_log_uniform;   /* 0: uniform in [lo,hi], 1: log-uniform      */
    float antiradiance_prob;    /* P(A = L), else A = 0                        */
    int with_boxes;             /* add the two rotated Cornell boxes          */
    int num_random_tris;        /* random triangles inside the unit cube      */
    float tri_area_lo, tri_area_hi; /* their areas, uniform in [lo, hi]       */
} arc_scene_desc;

/* Number of entries of the synthetic material table (index 0 is the unused
 * dummy entry, as in CMaterialBuffer.cpp:13-17). */
int arc_scene_num_materials(void);

/* Material table as 10 floats per entry: emissive[3], diffuse[3],
 * specular[3], exponent (the MATERIAL layout of Material.h:6-12). */
int arc_scene_materials(float* out_mats, int max_entries);

/* Camera position the G-buffer was cast from (looking along +z) and the scene
 * extent |bbMax - bbMin| used for the photon radius (Scene.h:75). */
int arc_scene_camera(const arc_scene_desc* d, float* cam_pos3, float* scene_extent);

/* Ray-cast the G-buffer.  albedo4 may be NULL.  Returns the number of pixels
 * whose primary ray missed every surface (0 for the shipped scenes), or a
 * negative value on invalid input. */
int arc_scene_gbuffer(const arc_scene_desc* d, float* pos4, float* nrm4, float* albedo4);

/* Sample d->num_avpls AVPLs area-uniformly on the scene surfaces.  The
 * incident direction is cosine-distributed into the surface (dot(n, w) < 0).
 * Returns 0 on success. */
int arc_scene_avpls(const arc_scene_desc* d, float* pos3, float* nrm3, float* dir3,
                    float* rad3, float* antirad3, int* material, int* bounce);

/* LBVH point set: n/2 points uniform in [0,1)^3 and n - n/2 points from an
 * equal-weight mixture of `num_gaussians` isotropic Gaussians with standard
 * deviation sigma and centres uniform in [0,1]^3 (not clamped). */
int arc_scene_points(uint64_t seed, int n, int num_gaussians, float sigma, float* pos3);

/* Unit normals drawn uniformly on the sphere (for SimpleBvh inputs). */
int arc_scene_normals(uint64_t seed, int n, float* nrm3);

/* The scene's triangles in the Triangle convention of the AVPL shooter
 * (front face normal = normalize(cross(p2 - p0, p1 - p0))); a two-sided
 * random triangle is written twice, once per face.  Returns the count
 * (at most max_tris written); arc_scene_num_triangles gives it up front. */
int arc_scene_num_triangles(const arc_scene_desc* d);
int arc_scene_triangles(const arc_scene_desc* d, float* p0, float* p1, float* p2, int* material, int max_tris);

/* A ceiling area light for the shooter: centre, front, up, width and height,
 * radiance, and its emissive MATERIAL (10 floats).  Any pointer may be NULL. */
int arc_scene_light(const arc_scene_desc* d, float* center3, float* front3, float* up3, float* size2,
                    float* radiance3, float* material10);

#ifdef __cplusplus
}
#endif

#endif /* ARC_SCENE_H_ */
