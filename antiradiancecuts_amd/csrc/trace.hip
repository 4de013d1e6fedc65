// Ray tracing of the triangle scene on gfx950: the AVPL shooter and the
// G-buffer producer that feed the gather.
//
// The reference shoots AVPLs on the CPU (AvplShooter::createPath,
// AvplShooter.h:77-142, OpenMP over paths) against a kd-tree
// (KdTreeAccelerator.h:17-59, Scene::IntersectRayScene, Scene.cpp:243-255),
// then copies the AVPLs to the GPU every frame (AvplsGpu, data_gpu.cu:5-52).
// Here the paths run on the GPU and the AVPLs land in HBM where the gather
// reads them.
//
//   scene      triangles in the reference's Triangle layout (p0, p1, p2,
//              normal = normalize(cross(p2 - p0, p1 - p0)), material), one
//              64-byte record per triangle in input order; the leaves are
//              triangle pieces (large triangles split into slabs along their
//              longest axis), and the tree over them is a binned-SAH build
//              on the host at scene creation (static geometry: built once,
//              traversed by every path and pixel);
//   intersect  Triangle::intersect in FRONT_FACE mode (Triangle.h): Moller-
//              Trumbore on unnormalised barycentrics with the reference's
//              |det| < EPS = 1e-4 rejection; the closest hit with t > 0
//              (KdTreeAccelerator.cpp:301-320); a hit on an emissive
//              triangle counts as no hit (Scene.cpp:247-252).  The arithmetic
//              follows the reference's operation order in FP32 without
//              contraction (this unit is built with -ffp-contract=off), so a
//              CPU restatement finds the same triangle and distance;
//   shoot      createPath per path: light sample (AreaLight::samplePos),
//              cosine direction (sampleCosCone, Sampler.h:13-35), bounces
//              with Russian roulette 0.8, phong_eval / phong_sample
//              (Brdf.h:9-82) and the EPS offsets.  Random numbers come from a
//              counter-based hash of (seed, path, draw) instead of per-thread
//              mt19937_64 streams; each path's AVPLs are written in order at
//              an offset from an exclusive scan of the per-path counts, so
//              the output is deterministic;
//   gbuffer    one primary ray per pixel (pinhole camera), the closest
//              front-facing triangle, position (x, y, z, 1) and normal
//              (n, material) in the G-buffer layout the gather reads.
#include <algorithm>
#include <cmath>
#include <vector>

#include "arc_internal.hpp"

using arc::f3;

struct arc_trace_scene {
    arc_ctx* ctx = nullptr;
    int ntri = 0;
    int nleaf = 0;                    // leaves: triangle pieces (a large triangle is split)
    arc::DevBuf tri;                  // 4 float4 per triangle, input order
    arc::DevBuf leaf_tri;             // per leaf: its triangle
    arc::DevBuf leaf_box;             // per leaf: box lo, hi (2 float4)
    arc::DevBuf nodes;                // the tree's internal nodes (BvhNode layout, boxes of the pieces)
    arc::DevBuf mats;                 // material table, 256 entries
    arc::DevBuf emissive;             // per material: emissive != 0
    int max_depth = 0;                // deepest internal node (root = 1)
    // shooter output
    arc::DevBuf state, active, stage, path_offset, scan_blocks, out;
    struct {
        f3 *pos, *nrm, *rad, *in_dir, *dir;
        float* pdf;
        int *bounce, *mat, *draws, *count;
    } ps{};
};

namespace {

constexpr float kEps = 1e-4f;         // EPS (Defines.h:13)
constexpr float kRr = 0.8f;           // rrProb (AvplShooter.h:78)
constexpr float kTwoPi = 6.28318530718f;          // TWO_PI (Defines.h:5)
constexpr float kOneOverTwoPi = 0.15915494309f;   // ONE_OVER_TWO_PI (Defines.h:7)
constexpr float kOneOverPi = 0.318309886f;
constexpr int kThreads = 256;
constexpr int kMaxPieces = 8;         // most leaves one triangle is split into
constexpr int kTraceStack = 48;       // per-lane stack of the closest-hit traversal (LDS)
constexpr int kScan = 1024;

struct V {
    float x, y, z;
};
__host__ __device__ inline V mk(float x, float y, float z) { return V{x, y, z}; }
__device__ inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V operator*(float s, V a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ inline V operator*(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
// glm::dot and glm::cross operation order
__device__ inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V cross(V a, V b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ inline float length(V a) { return sqrtf(dot(a, a)); }
__device__ inline V normalize(V a) { return (1.0f / length(a)) * a; }
__device__ inline V ld(const f3& a) { return mk(a.x, a.y, a.z); }
__device__ inline V ld(const arc_float3& a) { return mk(a.x, a.y, a.z); }

struct Tri {  // 4 float4 per triangle
    float4 a;  // p0, material index (bits)
    float4 b;  // e1 = p2 - p0 (FRONT_FACE first edge)
    float4 c;  // e2 = p1 - p0
    float4 d;  // unit normal, emissive flag
};

struct SceneDev {
    const float4* tri;               // triangles in input order, 4 float4 each
    const arc_bvh_node* nodes;       // n - 1 internal nodes
    const int* leaf_tri;             // leaf -> triangle
    int n;                           // leaves
    const arc_material* mats;        // 256 entries
    const int* emissive;             // 256 entries
};

// Triangle::intersect (Triangle.h), FRONT_FACE: e1 = p2 - p0, e2 = p1 - p0.
__device__ inline bool tri_hit(const Tri& T, V o, V dir, float det_eps, float& dist) {
    const V p0 = mk(T.a.x, T.a.y, T.a.z), e1 = mk(T.b.x, T.b.y, T.b.z), e2 = mk(T.c.x, T.c.y, T.c.z);
    const V p = cross(dir, e2);
    const float det = dot(e1, p);
    if (det < det_eps && det > -det_eps) return false;
    const V t = o - p0;
    const float u = dot(t, p);
    if (u < 0.0f || u > det) return false;
    const V q = cross(t, e1);
    const float v = dot(dir, q);
    if (v < 0.0f || u + v > det) return false;
    dist = 1.f / det * dot(e2, q);
    return true;
}

__device__ inline Tri load_tri(const float4* tri, int s) {
    const float4* r = tri + 4 * (size_t)s;
    return Tri{r[0], r[1], r[2], r[3]};
}

// Slab test; returns the entry distance or INFINITY when the box is missed
// or lies beyond t_max.
__device__ inline float box_enter(const arc_bvh_node& b, V o, V inv, float t_max) {
    const float tx0 = (b.bbMin.x - o.x) * inv.x, tx1 = (b.bbMax.x - o.x) * inv.x;
    const float ty0 = (b.bbMin.y - o.y) * inv.y, ty1 = (b.bbMax.y - o.y) * inv.y;
    const float tz0 = (b.bbMin.z - o.z) * inv.z, tz1 = (b.bbMax.z - o.z) * inv.z;
    const float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    const float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), t_max));
    return t0 <= t1 ? t0 : INFINITY;
}

__device__ inline float safe_inv(float w) {
    const float m = fmaxf(fabsf(w), 1e-30f);
    return 1.0f / copysignf(m, w);
}

// Closest front-facing hit with t > 0 (ties: the lower triangle index, as a
// loop over the triangles in order finds it).  stk: this lane's LDS stack,
// entry j at stk[j * kThreads].  Returns the triangle index or -1.  A
// triangle split into several leaves is tested once per leaf its ray
// reaches, with the same result each time.
__device__ int closest_hit(const SceneDev& S, V o, V dir, float det_eps, float& t_best, int* stk) {
    t_best = INFINITY;
    int best = -1;
    if (S.n == 1) {
        float t;
        const int ti = S.leaf_tri[0];
        if (tri_hit(load_tri(S.tri, ti), o, dir, det_eps, t) && t > 0.f) {
            t_best = t;
            best = ti;
        }
        return best;
    }
    const V inv = mk(safe_inv(dir.x), safe_inv(dir.y), safe_inv(dir.z));
    int sp = 0;
    int node = 0;  // internal node index
    for (;;) {
        const arc_bvh_node& b = S.nodes[node];
        int next[2];
        float tn[2];
        const int ch[2] = {b.left, b.right};
        int cnt = 0;
        for (int k = 0; k < 2; ++k) {
            const int c = ch[k];
            if (c >= 0) {  // leaf: test its triangle at once
                float t;
                const int ti = S.leaf_tri[c];
                if (tri_hit(load_tri(S.tri, ti), o, dir, det_eps, t) && t > 0.f &&
                    (t < t_best || (t == t_best && ti < best))) {
                    t_best = t;
                    best = ti;
                }
            } else {
                const float te = box_enter(S.nodes[-(c + 1)], o, inv, t_best);
                if (te < INFINITY) {
                    next[cnt] = -(c + 1);
                    tn[cnt] = te;
                    ++cnt;
                }
            }
        }
        if (cnt == 2) {
            const int nearer = tn[0] <= tn[1] ? 0 : 1;
            stk[(sp++) * kThreads] = next[1 - nearer];  // depth is checked at build: sp < kTraceStack
            node = next[nearer];
            continue;
        }
        if (cnt == 1) {
            node = next[0];
            continue;
        }
        // pop the next pending subtree that can still hold a closer hit
        bool found = false;
        while (sp > 0) {
            const int m = stk[(--sp) * kThreads];
            if (box_enter(S.nodes[m], o, inv, t_best) < INFINITY) {
                node = m;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

// Scene::IntersectRayScene: a hit on an emissive triangle is no hit.
__device__ inline int scene_hit(const SceneDev& S, V o, V dir, float& t, int* stk) {
    const int s = closest_hit(S, o, dir, kEps, t, stk);
    if (s < 0) return -1;
    const int mat = __float_as_int(S.tri[4 * (size_t)s].w);
    return S.emissive[mat & 255] ? -1 : s;
}

// ------------------------------------------------------------ sampling --
struct Rng {
    uint64_t seed, base;
    uint32_t k;
    __device__ inline float next() { return arc::hash_uniform(seed, base | (uint64_t)(k++)); }
};

// tangentSpace (Util.h:8-14): columns U, V, n.
__device__ inline void tangent_space(V n, V& U, V& W) {
    const V t = mk(n.y, n.z, -n.x);
    U = normalize(cross(t, n));
    W = normalize(cross(U, n));
}

// sampleCosCone (Sampler.h:13-35) with order `order`.
__device__ inline V sample_cos_cone(V n, float u1, float u2, float order, float& pdf) {
    V U, W;
    tangent_space(n, U, W);
    // cos(theta) = u1^(1 / (order + 1)); for the cosine lobe (order 1) the
    // square root, correctly rounded, and sin(theta) = sqrt(1 - cos^2) stand
    // for the reference's acos / pow / sin / cos (the same values up to
    // rounding, without acos's ill conditioning near theta = 0)
    float ct, st;
    if (order == 1.f) {
        ct = sqrtf(u1);
        st = sqrtf(fmaxf(0.f, 1.f - u1));
    } else {
        const float theta = acosf(powf(u1, 1.0f / (order + 1.0f)));
        ct = cosf(theta);
        st = sinf(theta);
    }
    const float phi = kTwoPi * u2;
    const V s = normalize(mk(st * cosf(phi), st * sinf(phi), ct));
    const float cos_theta = s.z;
    pdf = (order + 1.f) * kOneOverTwoPi * powf(cos_theta, order);
    return normalize(mk(U.x * s.x + W.x * s.y + n.x * s.z, U.y * s.x + W.y * s.y + n.y * s.z,
                        U.z * s.x + W.z * s.y + n.z * s.z));
}

__device__ inline float luminance(V v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
__device__ inline V reflect(V v, V n) { return normalize((2.f * dot(v, n)) * n - v); }

// phong_eval (Brdf.h:9-22)
__device__ inline V phong_eval(V wi, V wo, V n, const arc_material& m) {
    if (dot(wi, n) <= 0.f) return mk(0.f, 0.f, 0.f);
    if (dot(wo, n) <= 0.f) return mk(0.f, 0.f, 0.f);
    V res = mk(0.f, 0.f, 0.f);
    const V spec = ld(m.specular);
    if (length(spec) > 0.f) {
        const V r = reflect(wi, n);
        const float c = fmaxf(0.f, dot(r, wo));
        res = res + (kOneOverTwoPi * (m.exponent + 2.f) * powf(c, m.exponent)) * spec;
    }
    return res + kOneOverPi * ld(m.diffuse);
}

// phong_sample (Brdf.h:54-82): a sampled specular direction only decides
// whether the path ends; the returned direction is always the cosine sample
// about n.
__device__ inline V phong_sample(V wi, V n, const arc_material& m, float s0, float s1, float& pdf) {
    const bool has_diffuse = length(ld(m.diffuse)) > 0.f;
    const bool has_specular = length(ld(m.specular)) > 0.f;
    bool chose_specular = has_specular;
    if (has_diffuse && has_specular) {
        const float alpha = luminance(ld(m.specular)) / (luminance(ld(m.specular)) + luminance(ld(m.diffuse)));
        if (s0 <= alpha) {
            s0 /= alpha;
        } else {
            s0 = (s0 - alpha) / (1.f - alpha);
            chose_specular = false;
        }
    }
    if (chose_specular) {
        const V dir = sample_cos_cone(wi, s0, s1, m.exponent, pdf);
        if (dot(dir, n) <= 0.f) return mk(0.f, 0.f, 0.f);
    }
    return sample_cos_cone(n, s0, s1, 1.f, pdf);
}

struct Light {
    V center, front, s, u;  // quad centre, unit front, unit side axes
    float half_w, half_h, area;
    V radiance;
    int mat;
};

struct AvplOut {
    f3 *pos, *nrm, *dir, *rad, *anti;
    int *mat, *bounce;
};

// Wavefront form of createPath (AvplShooter.h:77-142): the state of every
// path lives in HBM between launches, one launch per bounce runs the paths
// still alive (no lane waits on another lane's longer path), and each
// launch stages the AVPLs it creates; a final scatter puts every path's
// AVPLs together in path order.  Draw order per path as in createPath:
// light position (2), first direction (2), then per bounce roulette (1) and
// the next direction (2).
struct PathState {
    f3 *pos, *nrm, *rad, *in_dir, *dir;
    float* pdf;
    int *bounce, *mat, *draws, *count;
};

struct Staged {  // AVPLs of all bounces, in the order the bounces created them
    AvplOut a;
    int* path;
};

__global__ __launch_bounds__(kThreads) void k_path_start(Light L, uint64_t seed, int paths, PathState P,
                                                         int* __restrict__ active) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= paths) return;
    Rng rng{seed, (uint64_t)i << 24, 0};
    // AreaLight::samplePos: glm::linearRand over [-1, 1]^2, x then y
    const float sx = -1.f + 2.f * rng.next();
    const float sy = -1.f + 2.f * rng.next();
    const V pos = L.center + (sx * L.half_w) * L.s + (sy * L.half_h) * L.u;
    const float pdf0 = 1.f / L.area;
    const V rad = (1.f / pdf0) * L.radiance;  // rad / pdf
    const float u1 = rng.next(), u2 = rng.next();
    float pdf;
    const V dir = sample_cos_cone(L.front, u1, u2, 1.f, pdf);
    P.pos[i] = f3{pos.x, pos.y, pos.z};
    P.nrm[i] = f3{L.front.x, L.front.y, L.front.z};
    P.rad[i] = f3{rad.x, rad.y, rad.z};
    P.in_dir[i] = f3{0.f, 0.f, 0.f};
    P.dir[i] = f3{dir.x, dir.y, dir.z};
    P.pdf[i] = pdf;
    P.bounce[i] = 0;
    P.mat[i] = L.mat;
    P.draws[i] = (int)rng.k;
    P.count[i] = 0;
    active[i] = i;
}

// One bounce of every live path: trace, create the AVPL, roulette, sample
// the next direction.  Survivors are appended to next_active.
__global__ __launch_bounds__(kThreads) void k_path_step(SceneDev S, uint64_t seed, int explicit_direct, PathState P,
                                                        const int* __restrict__ active, int n_active,
                                                        int* __restrict__ next_active, int* __restrict__ n_next,
                                                        Staged st, int stage_base) {
    __shared__ int stk[kTraceStack * kThreads];
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= n_active) return;
    const int i = active[j];
    const V nrm = ld(P.nrm[i]), direction = ld(P.dir[i]);
    const V org = ld(P.pos[i]) + kEps * nrm;
    float t;
    const int s = scene_hit(S, org, direction, t, stk + threadIdx.x);
    // every active path stages one AVPL slot; a miss leaves it unused (bounce 0)
    const int slot = stage_base + j;
    if (s < 0) {
        st.path[slot] = -1;
        return;
    }
    const V rad0 = ld(P.rad[i]);
    const float pdf0 = P.pdf[i];
    const float cos_theta = dot(nrm, direction);
    // getIncidentRadiance() * cos_theta / pdf, per component
    V contrib = mk(rad0.x * cos_theta / pdf0, rad0.y * cos_theta / pdf0, rad0.z * cos_theta / pdf0);
    const int bounce0 = P.bounce[i];
    if (bounce0 > 0) contrib = contrib * phong_eval(-1.f * ld(P.in_dir[i]), direction, nrm, S.mats[P.mat[i] & 255]);
    const float4* T = S.tri + 4 * (size_t)s;
    const V hit_n = mk(T[3].x, T[3].y, T[3].z);
    const V hit = org + t * direction;  // Intersection::setPosition(origin + dist * dir)
    const V pos = hit - kEps * hit_n;
    const int bounce = bounce0 + 1, mat = __float_as_int(T[0].w);
    V anti = contrib;
    if (bounce == 1 && explicit_direct) anti = mk(0.f, 0.f, 0.f);
    Rng rng{seed, (uint64_t)i << 24, (uint32_t)P.draws[i]};
    const bool stop = rng.next() > kRr;  // Russian roulette
    const V rad = stop ? mk(0.f, 0.f, 0.f) : (1.f / kRr) * contrib;
    st.a.pos[slot] = f3{pos.x, pos.y, pos.z};
    st.a.nrm[slot] = f3{hit_n.x, hit_n.y, hit_n.z};
    st.a.dir[slot] = f3{direction.x, direction.y, direction.z};
    st.a.rad[slot] = f3{rad.x, rad.y, rad.z};
    st.a.anti[slot] = f3{anti.x, anti.y, anti.z};
    st.a.mat[slot] = mat;
    st.a.bounce[slot] = bounce;
    st.path[slot] = i;
    P.count[i] += 1;
    if (stop) return;
    const float s0 = rng.next(), s1 = rng.next();
    float pdf;
    const V next = phong_sample(-1.f * direction, hit_n, S.mats[mat & 255], s0, s1, pdf);
    if (length(next) <= 0.f || pdf < 1e-4f) return;
    P.pos[i] = f3{pos.x, pos.y, pos.z};
    P.nrm[i] = f3{hit_n.x, hit_n.y, hit_n.z};
    P.rad[i] = f3{rad.x, rad.y, rad.z};
    P.in_dir[i] = f3{direction.x, direction.y, direction.z};
    P.dir[i] = f3{next.x, next.y, next.z};
    P.pdf[i] = pdf;
    P.bounce[i] = bounce;
    P.mat[i] = mat;
    P.draws[i] = (int)rng.k;
    next_active[atomicAdd(n_next, 1)] = i;
}

// Staged AVPL -> offset[path] + bounce - 1: each path's AVPLs together, in
// path order and bounce order.
__global__ __launch_bounds__(kThreads) void k_path_scatter(Staged st, int n_staged, const int* __restrict__ offsets,
                                                           AvplOut o) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n_staged) return;
    const int i = st.path[k];
    if (i < 0) return;
    const int dst = offsets[i] + st.a.bounce[k] - 1;
    o.pos[dst] = st.a.pos[k];
    o.nrm[dst] = st.a.nrm[k];
    o.dir[dst] = st.a.dir[k];
    o.rad[dst] = st.a.rad[k];
    o.anti[dst] = st.a.anti[k];
    o.mat[dst] = st.a.mat[k];
    o.bounce[dst] = st.a.bounce[k];
}

// Exclusive scan of n ints: per-block scan (kScan per block), block totals,
// then the totals' scan added back.  out[n] = total.
__global__ __launch_bounds__(kScan) void k_scan_local(const int* __restrict__ in, int n, int* __restrict__ out,
                                                      int* __restrict__ block_sums) {
    __shared__ int sh[kScan];
    const int i = blockIdx.x * kScan + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScan; off <<= 1) {
        const int a = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == kScan - 1) block_sums[blockIdx.x] = sh[kScan - 1];
}

__global__ __launch_bounds__(kScan) void k_scan_sums(int* __restrict__ sums, int nb) {
    __shared__ int sh[kScan];
    int run = 0;
    for (int base = 0; base < nb; base += kScan) {
        const int t = base + (int)threadIdx.x;
        const int v = t < nb ? sums[t] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kScan; off <<= 1) {
            const int a = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < nb) sums[t] = run + sh[threadIdx.x] - v;
        run += sh[kScan - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = run;
}

__global__ __launch_bounds__(kScan) void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ sums,
                                                    int nb) {
    const int i = blockIdx.x * kScan + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
    if (i == 0) out[n] = sums[nb];
}

// ---------------------------------------------------------- G-buffer --
struct Camera {
    V pos, right, up, fwd;  // right and up scaled by tan(fovy / 2) (and the aspect)
};

__global__ __launch_bounds__(kThreads) void k_gbuffer(SceneDev S, Camera cam, int W, int H,
                                                      float4* __restrict__ pos4, float4* __restrict__ nrm4) {
    __shared__ int stk[kTraceStack * kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= W * H) return;
    const int px = i % W, py = i / W;
    const float sx = 2.f * ((float)px + 0.5f) / (float)W - 1.f;
    const float sy = 1.f - 2.f * ((float)py + 0.5f) / (float)H;
    const V dir = normalize(cam.fwd + sx * cam.right + sy * cam.up);
    float t;
    // a rasteriser has no |det| floor: any front face counts
    const int s = closest_hit(S, cam.pos, dir, 0.f, t, stk + threadIdx.x);
    if (s < 0) {
        pos4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        nrm4[i] = make_float4(0.f, 0.f, -1.f, 1.f);
        return;
    }
    const float4* T = S.tri + 4 * (size_t)s;
    const V x = cam.pos + t * dir;
    pos4[i] = make_float4(x.x, x.y, x.z, 1.f);
    nrm4[i] = make_float4(T[3].x, T[3].y, T[3].z, (float)__float_as_int(T[0].w));
}

// --------------------------------------------------------------- build --
// Packs the triangles in sorted order and their boxes.
__global__ __launch_bounds__(kThreads) void k_pack_tris(const f3* __restrict__ p0, const f3* __restrict__ p1,
                                                        const f3* __restrict__ p2, const int* __restrict__ mat,
                                                        const int* __restrict__ emis, int n,
                                                        float4* __restrict__ tri) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int s = i;
    const V a = ld(p0[i]), b = ld(p1[i]), c = ld(p2[i]);
    V nn = cross(c - a, b - a);  // Triangle::calcNormal
    if (length(nn) > 0.f) nn = normalize(nn);
    const int m = mat[i];
    float4* o = tri + 4 * (size_t)s;
    o[0] = make_float4(a.x, a.y, a.z, __int_as_float(m));
    const V e1 = c - a, e2 = b - a;
    o[1] = make_float4(e1.x, e1.y, e1.z, 0.f);
    o[2] = make_float4(e2.x, e2.y, e2.z, 0.f);
    o[3] = make_float4(nn.x, nn.y, nn.z, (float)emis[m & 255]);
}


__global__ __launch_bounds__(kThreads) void k_zero_ints(int* __restrict__ p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kThreads) void k_intersect(SceneDev S, const f3* __restrict__ org,
                                                        const f3* __restrict__ dir, int n, float* __restrict__ t,
                                                        int* __restrict__ tri_index) {
    __shared__ int stk[kTraceStack * kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    float th;
    const int s = scene_hit(S, ld(org[i]), ld(dir[i]), th, stk + threadIdx.x);
    t[i] = s < 0 ? -1.f : th;
    tri_index[i] = s;
}

int blocks(int n) { return (n + kThreads - 1) / kThreads; }

SceneDev scene_dev(arc_trace_scene* sc) {
    return SceneDev{sc->tri.as<float4>(), sc->nodes.as<arc_bvh_node>(), sc->leaf_tri.as<int>(), sc->nleaf,
                    sc->mats.as<arc_material>(), sc->emissive.as<int>()};
}

// Binned SAH build over the leaves' boxes (host, once per scene): 16 bins
// of the centroids along the longest axis of their bounds, the split of
// least area x count on either side; a run of equal centroids splits at its
// middle.  Below depth kSahDepth the splits take the median instead, which
// bounds the tree's depth by kSahDepth + log2(leaves) for the traversal's
// stack.  Writes the leaf order (leaf slot -> box index) and the internal
// nodes in the library's BvhNode layout (child >= 0: leaf slot; < 0: internal
// -(k + 1)), with their boxes; node 0 is the root.
constexpr int kSahBins = 16;
constexpr int kSahDepth = 26;

void sah_build(const std::vector<float>& lo, const std::vector<float>& hi, std::vector<int>& order,
               std::vector<arc_bvh_node>& nodes) {
    const int n = (int)(lo.size() / 3);
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    nodes.assign(n > 1 ? n - 1 : 0, arc_bvh_node{});
    if (n < 2) return;
    std::vector<float> cen(3 * (size_t)n);
    for (size_t k = 0; k < cen.size(); ++k) cen[k] = 0.5f * (lo[k] + hi[k]);
    auto area = [](const float* l, const float* h) {
        const float x = h[0] - l[0], y = h[1] - l[1], z = h[2] - l[2];
        return x * y + y * z + z * x;
    };
    struct Task { int b, e, node, depth; };
    std::vector<Task> stack{{0, n, 0, 0}};
    int next = 1;
    while (!stack.empty()) {
        const Task T = stack.back();
        stack.pop_back();
        float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int j = T.b; j < T.e; ++j) {
            const int i = order[j];
            for (int d = 0; d < 3; ++d) {
                bl[d] = std::min(bl[d], lo[3 * i + d]);
                bh[d] = std::max(bh[d], hi[3 * i + d]);
                cl[d] = std::min(cl[d], cen[3 * i + d]);
                ch[d] = std::max(ch[d], cen[3 * i + d]);
            }
        }
        arc_bvh_node& nd = nodes[T.node];
        nd.bbMin = arc_float3{bl[0], bl[1], bl[2]};
        nd.bbMax = arc_float3{bh[0], bh[1], bh[2]};
        nd.visited = 0;
        int a = 0;
        for (int d = 1; d < 3; ++d)
            if (ch[d] - cl[d] > ch[a] - cl[a]) a = d;
        const int count = T.e - T.b;
        int mid = T.b + count / 2;
        bool binned = false;
        const float span = ch[a] - cl[a];
        if (span > 0.f && T.depth < kSahDepth) {
            int bc[kSahBins] = {};
            float bbl[kSahBins][3], bbh[kSahBins][3];
            for (int k = 0; k < kSahBins; ++k)
                for (int d = 0; d < 3; ++d) {
                    bbl[k][d] = INFINITY;
                    bbh[k][d] = -INFINITY;
                }
            const float scale = kSahBins / span;
            auto bin_of = [&](int i) {
                return std::min(kSahBins - 1, (int)((cen[3 * i + a] - cl[a]) * scale));
            };
            for (int j = T.b; j < T.e; ++j) {
                const int i = order[j], k = bin_of(i);
                ++bc[k];
                for (int d = 0; d < 3; ++d) {
                    bbl[k][d] = std::min(bbl[k][d], lo[3 * i + d]);
                    bbh[k][d] = std::max(bbh[k][d], hi[3 * i + d]);
                }
            }
            // sweep: areas of the left prefixes and right suffixes
            float ra[kSahBins];
            int rc[kSahBins];
            float sl[3] = {INFINITY, INFINITY, INFINITY}, sh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int cnt = 0;
            for (int k = kSahBins - 1; k > 0; --k) {
                cnt += bc[k];
                for (int d = 0; d < 3; ++d) {
                    sl[d] = std::min(sl[d], bbl[k][d]);
                    sh[d] = std::max(sh[d], bbh[k][d]);
                }
                ra[k] = cnt ? area(sl, sh) : 0.f;
                rc[k] = cnt;
            }
            float best = INFINITY;
            int split = -1;
            for (int d = 0; d < 3; ++d) {
                sl[d] = INFINITY;
                sh[d] = -INFINITY;
            }
            cnt = 0;
            for (int k = 0; k < kSahBins - 1; ++k) {
                cnt += bc[k];
                for (int d = 0; d < 3; ++d) {
                    sl[d] = std::min(sl[d], bbl[k][d]);
                    sh[d] = std::max(sh[d], bbh[k][d]);
                }
                if (cnt == 0 || rc[k + 1] == 0) continue;
                const float cost = (float)cnt * area(sl, sh) + (float)rc[k + 1] * ra[k + 1];
                if (cost < best) {
                    best = cost;
                    split = k;
                }
            }
            if (split >= 0) {  // both sides hold leaves by construction
                auto it = std::partition(order.begin() + T.b, order.begin() + T.e,
                                         [&](int i) { return bin_of(i) <= split; });
                mid = (int)(it - order.begin());
                binned = mid > T.b && mid < T.e;
            }
        }
        if (!binned) {  // median split (deep subtrees, runs of equal centroids)
            mid = T.b + count / 2;
            std::nth_element(order.begin() + T.b, order.begin() + mid, order.begin() + T.e,
                             [&](int x, int y) { return cen[3 * x + a] < cen[3 * y + a]; });
        }
        const int ranges[2][2] = {{T.b, mid}, {mid, T.e}};
        int child[2];
        for (int c = 0; c < 2; ++c) {
            if (ranges[c][1] - ranges[c][0] == 1) {
                child[c] = ranges[c][0];  // leaf slot
            } else {
                const int k = next++;
                child[c] = -(k + 1);
                stack.push_back({ranges[c][0], ranges[c][1], k, T.depth + 1});
            }
        }
        nd.left = child[0];
        nd.right = child[1];
    }
}

// Depth of the deepest internal node (host, from the node array).
int tree_depth(const std::vector<arc_bvh_node>& nodes) {
    if (nodes.empty()) return 0;
    int best = 0;
    std::vector<std::pair<int, int>> st{{0, 1}};
    while (!st.empty()) {
        auto [k, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        if (nodes[k].left < 0) st.push_back({-(nodes[k].left + 1), d + 1});
        if (nodes[k].right < 0) st.push_back({-(nodes[k].right + 1), d + 1});
    }
    return best;
}

}  // namespace

extern "C" {

int arc_trace_scene_create(arc_ctx* ctx, const arc_float3* p0, const arc_float3* p1, const arc_float3* p2,
                           const int* material, int ntri, const arc_material* materials, int num_materials,
                           int flags, arc_trace_scene** out) {
    if (!ctx || !out || ntri < 1 || !p0 || !p1 || !p2 || !material || !materials || num_materials < 1 ||
        num_materials > 256)
        return ARC_E_INVALID;
    *out = nullptr;
    arc_trace_scene* sc = new arc_trace_scene();
    sc->ctx = ctx;
    sc->ntri = ntri;
    auto fail = [&](int st) {
        arc_trace_scene_destroy(sc);
        return st;
    };
    // the triangles on the host (the tree is built there)
    const size_t v3 = sizeof(arc_float3) * (size_t)ntri;
    std::vector<arc_float3> h0(ntri), h1(ntri), h2(ntri);
    std::vector<int> hm(ntri);
    hipError_t e = hipSuccess;
    if (flags & ARC_INPUT_HOST) {
        std::copy(p0, p0 + ntri, h0.begin());
        std::copy(p1, p1 + ntri, h1.begin());
        std::copy(p2, p2 + ntri, h2.begin());
        std::copy(material, material + ntri, hm.begin());
    } else {
        e = hipMemcpy(h0.data(), p0, v3, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(h1.data(), p1, v3, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(h2.data(), p2, v3, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hm.data(), material, sizeof(int) * ntri, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(arc::hip_fail(ctx, e, "arc_trace_scene_create", __FILE__, __LINE__));
    }
    // Leaves are triangle pieces: a triangle larger than the median
    // triangle is split along its box's longest axis into up to kMaxPieces
    // slabs (early split clipping), each piece boxed by the triangle clipped
    // to its slab.  Rays leaving a surface start inside the boxes of every
    // large triangle around them and must open all of them; smaller boxes
    // mean fewer.  The clipping runs in double and the boxes are rounded
    // outward, so a piece's box always holds its part of the triangle.
    std::vector<float> plo, phi;
    std::vector<int> ptri;
    {
        std::vector<double> ext(ntri);
        for (int i = 0; i < ntri; ++i) {
            const double e0 = std::max({h0[i].x, h1[i].x, h2[i].x}) - std::min({h0[i].x, h1[i].x, h2[i].x});
            const double e1 = std::max({h0[i].y, h1[i].y, h2[i].y}) - std::min({h0[i].y, h1[i].y, h2[i].y});
            const double e2 = std::max({h0[i].z, h1[i].z, h2[i].z}) - std::min({h0[i].z, h1[i].z, h2[i].z});
            ext[i] = std::max({e0, e1, e2});
        }
        std::vector<double> sorted = ext;
        std::nth_element(sorted.begin(), sorted.begin() + ntri / 2, sorted.end());
        const double unit = sorted[ntri / 2];
        struct D3 { double v[3]; };
        auto clip = [](const std::vector<D3>& in, int a, double c, bool keep_ge) {
            std::vector<D3> out;
            for (size_t k = 0; k < in.size(); ++k) {
                const D3& p = in[k];
                const D3& q = in[(k + 1) % in.size()];
                const bool pin = keep_ge ? p.v[a] >= c : p.v[a] <= c;
                const bool qin = keep_ge ? q.v[a] >= c : q.v[a] <= c;
                if (pin) out.push_back(p);
                if (pin != qin) {
                    const double t = (c - p.v[a]) / (q.v[a] - p.v[a]);
                    D3 r;
                    for (int d = 0; d < 3; ++d) r.v[d] = p.v[d] + t * (q.v[d] - p.v[d]);
                    r.v[a] = c;
                    out.push_back(r);
                }
            }
            return out;
        };
        auto push = [&](int i, const double lo[3], const double hi[3], double pad) {
            for (int d = 0; d < 3; ++d) {
                plo.push_back(std::nextafter((float)(lo[d] - pad), -INFINITY));
                phi.push_back(std::nextafter((float)(hi[d] + pad), INFINITY));
            }
            ptri.push_back(i);
        };
        for (int i = 0; i < ntri; ++i) {
            const D3 v[3] = {{{h0[i].x, h0[i].y, h0[i].z}}, {{h1[i].x, h1[i].y, h1[i].z}},
                             {{h2[i].x, h2[i].y, h2[i].z}}};
            double blo[3], bhi[3];
            for (int d = 0; d < 3; ++d) {
                blo[d] = std::min({v[0].v[d], v[1].v[d], v[2].v[d]});
                bhi[d] = std::max({v[0].v[d], v[1].v[d], v[2].v[d]});
            }
            const double pad = 1e-6 * ext[i] + 1e-12;
            const int k = unit > 0.0 ? std::min(kMaxPieces, std::max(1, (int)std::ceil(ext[i] / unit))) : 1;
            if (k == 1) {
                push(i, blo, bhi, pad);
                continue;
            }
            int a = 0;
            for (int d = 1; d < 3; ++d)
                if (bhi[d] - blo[d] > bhi[a] - blo[a]) a = d;
            for (int j = 0; j < k; ++j) {
                const double lo_a = blo[a] + (bhi[a] - blo[a]) * j / k;
                const double hi_a = j == k - 1 ? bhi[a] : blo[a] + (bhi[a] - blo[a]) * (j + 1) / k;
                const std::vector<D3> poly = clip(clip({v[0], v[1], v[2]}, a, lo_a, true), a, hi_a, false);
                double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                for (const D3& q : poly)
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = std::min(lo[d], q.v[d]);
                        hi[d] = std::max(hi[d], q.v[d]);
                    }
                if (poly.empty())  // rounding only: keep the whole triangle box within the slab
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = blo[d];
                        hi[d] = bhi[d];
                    }
                lo[a] = lo_a;
                hi[a] = hi_a;
                push(i, lo, hi, pad);
            }
        }
    }
    const int np = (int)ptri.size();
    sc->nleaf = np;
    std::vector<arc_material> table(256);
    std::vector<int> emis(256, 0);
    for (int k = 0; k < num_materials; ++k) {
        table[k] = materials[k];
        const arc_float3 em = materials[k].emissive;
        emis[k] = (em.x * em.x + em.y * em.y + em.z * em.z) > 0.f ? 1 : 0;  // glm::length(emissive) > 0
    }
    int st = arc::ensure(ctx, sc->mats, sizeof(arc_material) * 256);
    if (st == ARC_OK) st = arc::ensure(ctx, sc->emissive, sizeof(int) * 256);
    if (st == ARC_OK) st = arc::ensure(ctx, sc->tri, sizeof(float4) * 4 * (size_t)ntri);
    if (st == ARC_OK) st = arc::ensure(ctx, sc->leaf_tri, sizeof(int) * (size_t)np);
    if (st == ARC_OK) st = arc::ensure(ctx, sc->leaf_box, sizeof(float4) * 2 * (size_t)np);
    arc::DevBuf staging;
    if (st == ARC_OK) st = arc::ensure(ctx, staging, 3 * v3 + sizeof(int) * (size_t)ntri);
    if (st != ARC_OK) {
        arc::release(staging);
        return fail(st);
    }
    char* base = staging.as<char>();
    f3* d0 = reinterpret_cast<f3*>(base);
    f3* d1 = reinterpret_cast<f3*>(base + v3);
    f3* d2 = reinterpret_cast<f3*>(base + 2 * v3);
    int* dm = reinterpret_cast<int*>(base + 3 * v3);
    e = hipMemcpy(d0, h0.data(), v3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d1, h1.data(), v3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d2, h2.data(), v3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dm, hm.data(), sizeof(int) * ntri, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sc->mats.p, table.data(), sizeof(arc_material) * 256, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sc->emissive.p, emis.data(), sizeof(int) * 256, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        arc::release(staging);
        return fail(arc::hip_fail(ctx, e, "arc_trace_scene_create", __FILE__, __LINE__));
    }
    // the tree: binned SAH over the pieces' boxes, built once on the host
    std::vector<int> order;
    std::vector<arc_bvh_node> hn;
    sah_build(plo, phi, order, hn);
    {
        std::vector<int> lt(np);
        std::vector<float4> lb(2 * (size_t)np);
        for (int k = 0; k < np; ++k) {
            const int r = order[k];
            lt[k] = ptri[r];
            lb[2 * (size_t)k] = make_float4(plo[3 * r], plo[3 * r + 1], plo[3 * r + 2], 0.f);
            lb[2 * (size_t)k + 1] = make_float4(phi[3 * r], phi[3 * r + 1], phi[3 * r + 2], 0.f);
        }
        e = hipMemcpy(sc->leaf_tri.p, lt.data(), sizeof(int) * (size_t)np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->leaf_box.p, lb.data(), sizeof(float4) * 2 * (size_t)np,
                                           hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            arc::release(staging);
            return fail(arc::hip_fail(ctx, e, "arc_trace_scene_create", __FILE__, __LINE__));
        }
    }
    hipLaunchKernelGGL(k_pack_tris, dim3(blocks(ntri)), dim3(kThreads), 0, ctx->stream, d0, d1, d2, dm,
                       sc->emissive.as<int>(), ntri, sc->tri.as<float4>());
    if (hipGetLastError() != hipSuccess) st = ARC_E_HIP;
    if (st == ARC_OK && np > 1) {
        const int nn = np - 1;
        st = arc::ensure(ctx, sc->nodes, sizeof(arc_bvh_node) * (size_t)nn);
        if (st == ARC_OK) {
            e = hipMemcpy(sc->nodes.p, hn.data(), sizeof(arc_bvh_node) * (size_t)nn, hipMemcpyHostToDevice);
            if (e != hipSuccess) st = arc::hip_fail(ctx, e, "arc_trace_scene_create", __FILE__, __LINE__);
        }
        sc->max_depth = tree_depth(hn);
        // the traversal keeps at most one pending subtree per level
        if (st == ARC_OK && sc->max_depth >= kTraceStack)
            st = arc::fail(ctx, ARC_E_UNSUPPORTED, "triangle tree deeper than the traversal stack");
    }
    if (st == ARC_OK) {
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = arc::hip_fail(ctx, e, "arc_trace_scene_create", __FILE__, __LINE__);
    }
    arc::release(staging);
    if (st != ARC_OK) return fail(st);
    *out = sc;
    return ARC_OK;
}

void arc_trace_scene_destroy(arc_trace_scene* sc) {
    if (!sc) return;
    if (sc->ctx && sc->ctx->stream) (void)hipStreamSynchronize(sc->ctx->stream);
    arc::release(sc->tri);
    arc::release(sc->leaf_tri);
    arc::release(sc->leaf_box);
    arc::release(sc->nodes);
    arc::release(sc->mats);
    arc::release(sc->emissive);
    arc::release(sc->state);
    arc::release(sc->active);
    arc::release(sc->stage);
    arc::release(sc->path_offset);
    arc::release(sc->scan_blocks);
    arc::release(sc->out);
    delete sc;
}

int arc_trace_intersect(arc_trace_scene* sc, const arc_float3* org, const arc_float3* dir, int n, float* t,
                        int* tri_index) {
    if (!sc || n < 0 || (n > 0 && (!org || !dir || !t || !tri_index))) return ARC_E_INVALID;
    if (n == 0) return ARC_OK;
    arc_ctx* ctx = sc->ctx;
    hipLaunchKernelGGL(k_intersect, dim3(blocks(n)), dim3(kThreads), 0, ctx->stream, scene_dev(sc),
                       reinterpret_cast<const f3*>(org), reinterpret_cast<const f3*>(dir), n, t, tri_index);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int arc_trace_shoot(arc_trace_scene* sc, const arc_area_light* light, int num_paths, uint64_t seed, int flags,
                    arc_avpls* out, int* num_indirect) {
    if (!sc || !light || !out || !num_indirect || num_paths < 0) return ARC_E_INVALID;
    arc_ctx* ctx = sc->ctx;
    *num_indirect = 0;
    // light frame (AreaLight::updateWorldTransform: lookAt towards -front)
    const float fl = std::sqrt(light->front.x * light->front.x + light->front.y * light->front.y +
                               light->front.z * light->front.z);
    if (!(fl > 0.f) || !(light->width > 0.f) || !(light->height > 0.f) || light->materialIndex < 0 ||
        light->materialIndex > 255)
        return arc::fail(ctx, ARC_E_INVALID, "area light");
    Light L;
    L.center = mk(light->center.x, light->center.y, light->center.z);
    L.front = mk(light->front.x / fl, light->front.y / fl, light->front.z / fl);
    {
        const float fx = -L.front.x, fy = -L.front.y, fz = -L.front.z;  // lookAt forward
        const float ux = light->up.x, uy = light->up.y, uz = light->up.z;
        float sx = fy * uz - fz * uy, sy = fz * ux - fx * uz, sz = fx * uy - fy * ux;  // cross(f, up)
        const float sl = std::sqrt(sx * sx + sy * sy + sz * sz);
        if (!(sl > 0.f)) return arc::fail(ctx, ARC_E_INVALID, "area light: up parallel to front");
        sx /= sl; sy /= sl; sz /= sl;
        L.s = mk(sx, sy, sz);
        L.u = mk(sy * fz - sz * fy, sz * fx - sx * fz, sx * fy - sy * fx);  // cross(s, f)
    }
    L.half_w = 0.5f * light->width;
    L.half_h = 0.5f * light->height;
    L.area = light->width * light->height;
    L.radiance = mk(light->radiance.x, light->radiance.y, light->radiance.z);
    L.mat = light->materialIndex;
    const int explicit_direct = (flags & ARC_SHOOT_NO_EXPLICIT_DIRECT) ? 0 : 1;
    if (num_paths == 0) return ARC_OK;
    const size_t np = (size_t)num_paths;
    // path state, 80 bytes per path
    ARC_TRY(arc::ensure(ctx, sc->state, np * (5 * sizeof(f3) + 5 * sizeof(int))));
    {
        char* b = sc->state.as<char>();
        f3* v[5];
        for (int k = 0; k < 5; ++k) v[k] = reinterpret_cast<f3*>(b + (size_t)k * np * sizeof(f3));
        int* w = reinterpret_cast<int*>(b + 5 * np * sizeof(f3));
        sc->ps = {v[0], v[1], v[2], v[3], v[4], reinterpret_cast<float*>(w), w + np, w + 2 * np,
                               w + 3 * np, w + 4 * np};
    }
    const PathState P{sc->ps.pos, sc->ps.nrm, sc->ps.rad, sc->ps.in_dir, sc->ps.dir, sc->ps.pdf, sc->ps.bounce,
                      sc->ps.mat, sc->ps.draws, sc->ps.count};
    ARC_TRY(arc::ensure(ctx, sc->active, sizeof(int) * (2 * np + 2)));
    int* act[2] = {sc->active.as<int>(), sc->active.as<int>() + np};
    int* n_next = sc->active.as<int>() + 2 * np;
    const SceneDev S = scene_dev(sc);
    hipLaunchKernelGGL(k_path_start, dim3(blocks(num_paths)), dim3(kThreads), 0, ctx->stream, L, seed, num_paths, P,
                       act[0]);
    ARC_LAUNCHED(ctx);
    // one launch per bounce over the live paths
    size_t cap = 0, staged = 0;
    Staged st{};
    auto stage_at = [&](arc::DevBuf& buf, size_t c) {
        char* b = buf.as<char>();
        Staged r;
        r.a.pos = reinterpret_cast<f3*>(b);
        r.a.nrm = r.a.pos + c;
        r.a.dir = r.a.nrm + c;
        r.a.rad = r.a.dir + c;
        r.a.anti = r.a.rad + c;
        r.a.mat = reinterpret_cast<int*>(r.a.anti + c);
        r.a.bounce = r.a.mat + c;
        r.path = r.a.bounce + c;
        return r;
    };
    const size_t slot_bytes = 5 * sizeof(f3) + 3 * sizeof(int);
    if (sc->stage.p) {  // reuse the last call's staging area
        cap = sc->stage.bytes / slot_bytes;
        st = stage_at(sc->stage, cap);
    }
    int n_active = num_paths, cur = 0;
    for (int step = 0; n_active > 0; ++step) {
        if (step > 4096) return arc::fail(ctx, ARC_E_STATE, "path tracing did not terminate");
        if (staged + (size_t)n_active > cap) {  // grow the staging area, keeping what it holds
            const size_t ncap = std::max(staged + (size_t)n_active, std::max((size_t)8 * np, 2 * cap));
            arc::DevBuf nb;
            ARC_TRY(arc::ensure(ctx, nb, ncap * slot_bytes));
            const Staged ns = stage_at(nb, ncap);
            if (staged > 0) {
                const void* src[8] = {st.a.pos, st.a.nrm, st.a.dir, st.a.rad, st.a.anti, st.a.mat, st.a.bounce, st.path};
                void* dst[8] = {ns.a.pos, ns.a.nrm, ns.a.dir, ns.a.rad, ns.a.anti, ns.a.mat, ns.a.bounce, ns.path};
                const size_t sz[8] = {sizeof(f3), sizeof(f3), sizeof(f3), sizeof(f3), sizeof(f3), 4, 4, 4};
                for (int k = 0; k < 8; ++k)
                    ARC_HIP(ctx, hipMemcpyAsync(dst[k], src[k], staged * sz[k], hipMemcpyDeviceToDevice, ctx->stream));
            }
            ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            arc::release(sc->stage);
            sc->stage = nb;
            cap = ncap;
            st = ns;
        }
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(kThreads), 0, ctx->stream, n_next, 1);
        hipLaunchKernelGGL(k_path_step, dim3(blocks(n_active)), dim3(kThreads), 0, ctx->stream, S, seed,
                           explicit_direct, P, act[cur], n_active, act[1 - cur], n_next, st, (int)staged);
        ARC_LAUNCHED(ctx);
        staged += (size_t)n_active;
        ARC_HIP(ctx, hipMemcpyAsync(&n_active, n_next, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        cur = 1 - cur;
    }
    // per-path counts -> offsets -> AVPLs in path order
    const int nb = (num_paths + kScan - 1) / kScan;
    ARC_TRY(arc::ensure(ctx, sc->path_offset, sizeof(int) * (np + 1)));
    ARC_TRY(arc::ensure(ctx, sc->scan_blocks, sizeof(int) * ((size_t)nb + 1)));
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScan), 0, ctx->stream, P.count, num_paths,
                       sc->path_offset.as<int>(), sc->scan_blocks.as<int>());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScan), 0, ctx->stream, sc->scan_blocks.as<int>(), nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScan), 0, ctx->stream, sc->path_offset.as<int>(), num_paths,
                       sc->scan_blocks.as<int>(), nb);
    ARC_LAUNCHED(ctx);
    int total = 0;
    ARC_HIP(ctx, hipMemcpyAsync(&total, sc->path_offset.as<int>() + num_paths, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t v3 = sizeof(f3) * (size_t)std::max(total, 1), vi = sizeof(int) * (size_t)std::max(total, 1);
    ARC_TRY(arc::ensure(ctx, sc->out, 5 * v3 + 2 * vi));
    char* base = sc->out.as<char>();
    AvplOut o{reinterpret_cast<f3*>(base), reinterpret_cast<f3*>(base + v3), reinterpret_cast<f3*>(base + 2 * v3),
              reinterpret_cast<f3*>(base + 3 * v3), reinterpret_cast<f3*>(base + 4 * v3),
              reinterpret_cast<int*>(base + 5 * v3), reinterpret_cast<int*>(base + 5 * v3 + vi)};
    if (staged > 0) {
        hipLaunchKernelGGL(k_path_scatter, dim3(blocks((int)staged)), dim3(kThreads), 0, ctx->stream, st, (int)staged,
                           sc->path_offset.as<int>(), o);
        ARC_LAUNCHED(ctx);
    }
    out->position = reinterpret_cast<const arc_float3*>(o.pos);
    out->normal = reinterpret_cast<const arc_float3*>(o.nrm);
    out->incDirection = reinterpret_cast<const arc_float3*>(o.dir);
    out->incRadiance = reinterpret_cast<const arc_float3*>(o.rad);
    out->antiradiance = reinterpret_cast<const arc_float3*>(o.anti);
    out->materialIndex = o.mat;
    out->bounce = o.bounce;
    out->numAvpls = total;
    *num_indirect = total;
    return ARC_OK;
}

int arc_trace_gbuffer(arc_trace_scene* sc, arc_float3 camPos, arc_float3 forward, arc_float3 up, float fovy,
                      int width, int height, float* pos4, float* nrm4) {
    if (!sc || width <= 0 || height <= 0 || !pos4 || !nrm4 || !(fovy > 0.f) || !(fovy < 3.14159f))
        return ARC_E_INVALID;
    arc_ctx* ctx = sc->ctx;
    // camera basis: R = normalize(up x fwd), U = fwd x R
    const float fl = std::sqrt(forward.x * forward.x + forward.y * forward.y + forward.z * forward.z);
    if (!(fl > 0.f)) return ARC_E_INVALID;
    const float fx = forward.x / fl, fy = forward.y / fl, fz = forward.z / fl;
    float rx = up.y * fz - up.z * fy, ry = up.z * fx - up.x * fz, rz = up.x * fy - up.y * fx;
    const float rl = std::sqrt(rx * rx + ry * ry + rz * rz);
    if (!(rl > 0.f)) return ARC_E_INVALID;
    rx /= rl; ry /= rl; rz /= rl;
    const float ux = fy * rz - fz * ry, uy = fz * rx - fx * rz, uz = fx * ry - fy * rx;
    const float th = std::tan(0.5f * fovy), aspect = (float)width / (float)height;
    Camera cam{mk(camPos.x, camPos.y, camPos.z), mk(rx * th * aspect, ry * th * aspect, rz * th * aspect),
               mk(ux * th, uy * th, uz * th), mk(fx, fy, fz)};
    const int npix = width * height;
    hipLaunchKernelGGL(k_gbuffer, dim3(blocks(npix)), dim3(kThreads), 0, ctx->stream, scene_dev(sc), cam, width,
                       height, reinterpret_cast<float4*>(pos4), reinterpret_cast<float4*>(nrm4));
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int arc_trace_scene_info(arc_trace_scene* sc, int* num_triangles, int* tree_depth) {
    if (!sc) return ARC_E_INVALID;
    if (num_triangles) *num_triangles = sc->ntri;
    if (tree_depth) *tree_depth = sc->max_depth;
    return ARC_OK;
}

}  // extern "C"
