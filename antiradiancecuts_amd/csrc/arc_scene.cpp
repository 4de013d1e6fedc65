// Seeded synthetic scenes: a Cornell-style unit box (optionally with the two
// rotated boxes and a cloud of random triangles), ray-cast into the G-buffer
// layout of Shaders/CreateGBuffer.frag:14-15, AVPLs sampled on its surfaces
// with the fields of Avpl (AVPL.h:12-26), and the LBVH point sets.
//
// Host-only C++ (g++ -O2 -fopenmp).  Deterministic: each pixel, AVPL and point
// draws from its own SplitMix64 stream keyed by (seed, stream id, index).

#include "arc_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- RNG ----
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream, uint64_t index)
        : s(mix64(seed * 0x9e3779b97f4a7c15ULL + mix64(stream + 0x632be59bd9b4e019ULL) + index)) {}
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ULL;
        return mix64(s);
    }
    double ud() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

enum Stream : uint64_t {
    S_ALBEDO = 1, S_AVPL = 2, S_TRIS = 3, S_POINTS = 4, S_CENTRES = 5, S_NORMALS = 6
};

// ------------------------------------------------------------ vectors ----
struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 unit(V3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

void onb(V3 n, V3& t, V3& b) {
    V3 a = std::fabs(n.x) > 0.9 ? V3{0, 1, 0} : V3{1, 0, 0};
    t = unit(cross(a, n));
    b = cross(n, t);
}

V3 uniform_sphere(Rng& r) {
    double z = 1.0 - 2.0 * r.ud();
    double phi = 2.0 * M_PI * r.ud();
    double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    return {s * std::cos(phi), s * std::sin(phi), z};
}

// ------------------------------------------------------------- scene -----
struct Tri {
    V3 v0, v1, v2;
    V3 n;          // unit geometric normal
    int material;
    int two_sided; // random triangles may be seen / lit from either side
};

// Material table: 0 is the unused dummy entry, 1..5 are Lambertian.
const float kMaterials[6][10] = {
    {0, 0, 0, 0.0f, 0.0f, 0.0f, 0, 0, 0, 0},
    {0, 0, 0, 0.50f, 0.50f, 0.50f, 0, 0, 0, 0},   // white walls
    {0, 0, 0, 0.63f, 0.07f, 0.06f, 0, 0, 0, 0},   // red wall (x = 0)
    {0, 0, 0, 0.14f, 0.45f, 0.09f, 0, 0, 0, 0},   // green wall (x = 1)
    {0, 0, 0, 0.55f, 0.55f, 0.50f, 0, 0, 0, 0},   // boxes
    {0, 0, 0, 0.35f, 0.40f, 0.70f, 0, 0, 0, 0},   // blue-ish clutter
};
constexpr int kNumMaterials = 6;

void add_quad(std::vector<Tri>& tris, V3 a, V3 b, V3 c, V3 d, V3 n, int mat) {
    tris.push_back({a, b, c, n, mat, 0});
    tris.push_back({a, c, d, n, mat, 0});
}

// Upright box given its footprint (4 corners in x/z, counter-clockwise or
// clockwise) and height; top and four sides, normals pointing outwards.
void add_box(std::vector<Tri>& tris, const double fx[4], const double fz[4], double h, int mat) {
    double cx = 0.25 * (fx[0] + fx[1] + fx[2] + fx[3]);
    double cz = 0.25 * (fz[0] + fz[1] + fz[2] + fz[3]);
    V3 top[4], bot[4];
    for (int i = 0; i < 4; ++i) {
        top[i] = {fx[i], h, fz[i]};
        bot[i] = {fx[i], 0.0, fz[i]};
    }
    add_quad(tris, top[0], top[1], top[2], top[3], {0, 1, 0}, mat);
    for (int i = 0; i < 4; ++i) {
        int j = (i + 1) & 3;
        V3 n = unit(cross(top[j] - bot[i], bot[j] - bot[i]));
        V3 mid = 0.5 * (bot[i] + bot[j]);
        if (dot(n, mid - V3{cx, 0, cz}) < 0) n = -1.0 * n;
        add_quad(tris, bot[i], bot[j], top[j], top[i], n, mat);
    }
}

std::vector<Tri> build_scene(const arc_scene_desc& d) {
    std::vector<Tri> tris;
    // Unit box, open towards -z, normals facing inwards.
    add_quad(tris, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, 1);    // floor
    add_quad(tris, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, {0, -1, 0}, 1);   // ceiling
    add_quad(tris, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {0, 0, -1}, 1);   // back
    add_quad(tris, {0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, 2);    // left
    add_quad(tris, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {-1, 0, 0}, 3);   // right
    if (d.with_boxes) {
        // Classic Cornell box footprints (mm), scaled into the unit cube.
        const double sx = 1.0 / 556.0, sy = 1.0 / 548.8, sz = 1.0 / 559.2;
        const double s_x[4] = {130 * sx, 82 * sx, 240 * sx, 290 * sx};
        const double s_z[4] = {65 * sz, 225 * sz, 272 * sz, 114 * sz};
        const double t_x[4] = {423 * sx, 265 * sx, 314 * sx, 472 * sx};
        const double t_z[4] = {247 * sz, 296 * sz, 456 * sz, 406 * sz};
        add_box(tris, s_x, s_z, 165 * sy, 4);
        add_box(tris, t_x, t_z, 330 * sy, 4);
    }
    const int nr = std::max(0, d.num_random_tris);
    tris.reserve(tris.size() + nr);
    for (int i = 0; i < nr; ++i) {
        Rng r(d.seed, S_TRIS, (uint64_t)i);
        V3 c = {r.ud(), r.ud(), r.ud()};
        V3 n = uniform_sphere(r);
        V3 t, b;
        onb(n, t, b);
        double area = d.tri_area_lo + (d.tri_area_hi - d.tri_area_lo) * r.ud();
        // equilateral triangle: area = 3*sqrt(3)/4 * R^2
        double R = std::sqrt(area / (0.75 * std::sqrt(3.0)));
        double th = 2.0 * M_PI * r.ud();
        V3 v[3];
        for (int k = 0; k < 3; ++k) {
            double a = th + k * (2.0 * M_PI / 3.0);
            v[k] = c + (R * std::cos(a)) * t + (R * std::sin(a)) * b;
        }
        int mat = 1 + (int)(r.ud() * 5.0);
        if (mat > 5) mat = 5;
        tris.push_back({v[0], v[1], v[2], unit(cross(v[1] - v[0], v[2] - v[0])), mat, 1});
    }
    return tris;
}

double tri_area(const Tri& t) {
    V3 c = cross(t.v1 - t.v0, t.v2 - t.v0);
    return 0.5 * std::sqrt(dot(c, c));
}

// ------------------------------------------------------- uniform grid ----
struct Grid {
    int g;
    V3 lo;
    double cell;
    std::vector<int> start;  // g^3 + 1
    std::vector<int> items;

    int idx(int x, int y, int z) const { return (z * g + y) * g + x; }
};

bool plane_hits_box(const Tri& t, V3 bmin, V3 bmax) {
    V3 c = 0.5 * (bmin + bmax);
    V3 e = 0.5 * (bmax - bmin);
    double r = e.x * std::fabs(t.n.x) + e.y * std::fabs(t.n.y) + e.z * std::fabs(t.n.z);
    double s = dot(t.n, c - t.v0);
    return std::fabs(s) <= r * (1.0 + 1e-9) + 1e-12;
}

Grid build_grid(const std::vector<Tri>& tris) {
    Grid G;
    const size_t n = tris.size();
    G.g = n < 100 ? 4 : (n < 20000 ? 24 : 64);
    G.lo = {-0.15, -0.15, -0.15};
    G.cell = 1.3 / G.g;
    const int g = G.g;
    std::vector<int> count((size_t)g * g * g + 1, 0);
    auto cell_range = [&](const Tri& t, int lo[3], int hi[3]) {
        double mn[3] = {std::min({t.v0.x, t.v1.x, t.v2.x}), std::min({t.v0.y, t.v1.y, t.v2.y}),
                        std::min({t.v0.z, t.v1.z, t.v2.z})};
        double mx[3] = {std::max({t.v0.x, t.v1.x, t.v2.x}), std::max({t.v0.y, t.v1.y, t.v2.y}),
                        std::max({t.v0.z, t.v1.z, t.v2.z})};
        double o[3] = {G.lo.x, G.lo.y, G.lo.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::max(0, std::min(g - 1, (int)std::floor((mn[k] - o[k]) / G.cell - 1e-6)));
            hi[k] = std::max(0, std::min(g - 1, (int)std::floor((mx[k] - o[k]) / G.cell + 1e-6)));
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            G.start.assign(count.size(), 0);
            for (size_t i = 0; i + 1 < count.size(); ++i) G.start[i + 1] = G.start[i] + count[i];
            G.items.resize(G.start.back());
            std::fill(count.begin(), count.end(), 0);
        }
        for (size_t ti = 0; ti < n; ++ti) {
            int lo[3], hi[3];
            cell_range(tris[ti], lo, hi);
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        V3 bmin = {G.lo.x + x * G.cell, G.lo.y + y * G.cell, G.lo.z + z * G.cell};
                        V3 bmax = {bmin.x + G.cell, bmin.y + G.cell, bmin.z + G.cell};
                        if (!plane_hits_box(tris[ti], bmin, bmax)) continue;
                        int c = G.idx(x, y, z);
                        if (pass == 1) G.items[G.start[c] + count[c]] = (int)ti;
                        count[c]++;
                    }
        }
    }
    return G;
}

// Moller-Trumbore, double precision; returns t > tmin or -1.
double hit_tri(const Tri& t, V3 o, V3 dir) {
    V3 e1 = t.v1 - t.v0, e2 = t.v2 - t.v0;
    V3 p = cross(dir, e2);
    double det = dot(e1, p);
    if (std::fabs(det) < 1e-18) return -1.0;
    double inv = 1.0 / det;
    V3 s = o - t.v0;
    // a small tolerance closes the gap rounding opens along shared edges
    const double tol = 1e-9;
    double u = dot(s, p) * inv;
    if (u < -tol || u > 1.0 + tol) return -1.0;
    V3 q = cross(s, e1);
    double v = dot(dir, q) * inv;
    if (v < -tol || u + v > 1.0 + tol) return -1.0;
    double tt = dot(e2, q) * inv;
    return tt > 1e-9 ? tt : -1.0;
}

// Grid DDA; returns triangle index or -1.
int cast(const Grid& G, const std::vector<Tri>& tris, V3 o, V3 dir, double& t_hit) {
    const int g = G.g;
    const double o3[3] = {o.x, o.y, o.z}, d3[3] = {dir.x, dir.y, dir.z};
    const double lo3[3] = {G.lo.x, G.lo.y, G.lo.z};
    double t0 = 0.0, t1 = 1e30;
    for (int k = 0; k < 3; ++k) {
        double hi = lo3[k] + g * G.cell;
        if (std::fabs(d3[k]) < 1e-300) {
            if (o3[k] < lo3[k] || o3[k] > hi) return -1;
            continue;
        }
        double ta = (lo3[k] - o3[k]) / d3[k], tb = (hi - o3[k]) / d3[k];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
    }
    if (t0 > t1) return -1;
    int c[3], step[3];
    double tmax[3], tdelta[3];
    for (int k = 0; k < 3; ++k) {
        double p = o3[k] + t0 * d3[k];
        c[k] = std::max(0, std::min(g - 1, (int)std::floor((p - lo3[k]) / G.cell)));
        if (d3[k] > 0) {
            step[k] = 1;
            tmax[k] = (lo3[k] + (c[k] + 1) * G.cell - o3[k]) / d3[k];
            tdelta[k] = G.cell / d3[k];
        } else if (d3[k] < 0) {
            step[k] = -1;
            tmax[k] = (lo3[k] + c[k] * G.cell - o3[k]) / d3[k];
            tdelta[k] = -G.cell / d3[k];
        } else {
            step[k] = 0;
            tmax[k] = 1e300;
            tdelta[k] = 1e300;
        }
    }
    int best = -1;
    double best_t = 1e300;
    while (true) {
        int ci = G.idx(c[0], c[1], c[2]);
        for (int j = G.start[ci]; j < G.start[ci + 1]; ++j) {
            int ti = G.items[j];
            double tt = hit_tri(tris[ti], o, dir);
            if (tt > 0 && (tt < best_t || (tt == best_t && ti < best))) {
                best_t = tt;
                best = ti;
            }
        }
        int k = (tmax[0] < tmax[1]) ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
        if (best >= 0 && best_t <= tmax[k] + 1e-12) break;
        c[k] += step[k];
        if (c[k] < 0 || c[k] >= g) break;
        tmax[k] += tdelta[k];
    }
    t_hit = best_t;
    return best;
}

void camera_of(const arc_scene_desc& d, V3& cam, double& tan_half) {
    const double aspect = double(d.width) / double(d.height);
    tan_half = std::tan(0.5 * (M_PI / 4.0));  // 45 degree vertical field of view
    // keep the frustum inside the open face of the box at z = 0
    const double z0 = 0.49 / (tan_half * std::max(aspect, 1.0));
    cam = {0.5, 0.5, -z0};
}

bool valid(const arc_scene_desc* d) {
    return d && d->width > 0 && d->height > 0 && d->num_avpls >= 0 && d->num_random_tris >= 0;
}

}  // namespace

extern "C" {

int arc_scene_num_materials(void) { return kNumMaterials; }

int arc_scene_materials(float* out, int max_entries) {
    if (!out || max_entries < kNumMaterials) return -1;
    std::memcpy(out, kMaterials, sizeof(kMaterials));
    return kNumMaterials;
}

int arc_scene_camera(const arc_scene_desc* d, float* cam_pos3, float* scene_extent) {
    if (!valid(d)) return -1;
    V3 cam;
    double th;
    camera_of(*d, cam, th);
    cam_pos3[0] = (float)cam.x;
    cam_pos3[1] = (float)cam.y;
    cam_pos3[2] = (float)cam.z;
    // Scene bounding box is the unit cube (random triangles are centred in it;
    // the extent follows Scene::getSceneExtent = |bbMax - bbMin|).
    *scene_extent = (float)std::sqrt(3.0);
    return 0;
}

int arc_scene_gbuffer(const arc_scene_desc* d, float* pos4, float* nrm4, float* albedo4) {
    if (!valid(d) || !pos4 || !nrm4) return -1;
    const std::vector<Tri> tris = build_scene(*d);
    const Grid G = build_grid(tris);
    V3 cam;
    double tan_half;
    camera_of(*d, cam, tan_half);
    const int W = d->width, H = d->height;
    const double aspect = double(W) / double(H);
    int misses = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : misses)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            const size_t p = (size_t)py * W + px;
            V3 dir = unit(V3{(2.0 * (px + 0.5) / W - 1.0) * aspect * tan_half,
                             (1.0 - 2.0 * (py + 0.5) / H) * tan_half, 1.0});
            double t;
            int ti = cast(G, tris, cam, dir, t);
            if (ti < 0) {
                misses++;
                for (int k = 0; k < 4; ++k) pos4[4 * p + k] = 0.f;
                nrm4[4 * p + 0] = 0.f;
                nrm4[4 * p + 1] = 0.f;
                nrm4[4 * p + 2] = -1.f;
                nrm4[4 * p + 3] = 1.f;
            } else {
                V3 x = cam + t * dir;
                V3 n = tris[ti].n;
                if (dot(n, dir) > 0) n = -1.0 * n;  // face the camera
                pos4[4 * p + 0] = (float)x.x;
                pos4[4 * p + 1] = (float)x.y;
                pos4[4 * p + 2] = (float)x.z;
                pos4[4 * p + 3] = 1.f;
                nrm4[4 * p + 0] = (float)n.x;
                nrm4[4 * p + 1] = (float)n.y;
                nrm4[4 * p + 2] = (float)n.z;
                nrm4[4 * p + 3] = (float)tris[ti].material;
            }
            if (albedo4) {
                Rng r(d->seed, S_ALBEDO, p);
                for (int k = 0; k < 3; ++k) albedo4[4 * p + k] = (float)(0.2 + 0.6 * r.ud());
                albedo4[4 * p + 3] = 0.f;
            }
        }
    }
    return misses;
}

int arc_scene_avpls(const arc_scene_desc* d, float* pos3, float* nrm3, float* dir3,
                    float* rad3, float* antirad3, int* material, int* bounce) {
    if (!valid(d) || !pos3 || !nrm3 || !dir3 || !rad3 || !antirad3 || !material || !bounce)
        return -1;
    if (d->max_bounce < d->min_bounce || d->radiance_hi < d->radiance_lo) return -1;
    if (d->radiance_log_uniform && d->radiance_lo <= 0.f) return -1;
    const std::vector<Tri> tris = build_scene(*d);
    std::vector<double> cdf(tris.size() + 1, 0.0);
    for (size_t i = 0; i < tris.size(); ++i) cdf[i + 1] = cdf[i] + tri_area(tris[i]);
    const double total = cdf.back();
    const int N = d->num_avpls;
    const double llo = d->radiance_log_uniform ? std::log((double)d->radiance_lo) : 0.0;
    const double lhi = d->radiance_log_uniform ? std::log((double)d->radiance_hi) : 0.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        Rng r(d->seed, S_AVPL, (uint64_t)i);
        const double u = r.ud() * total;
        size_t ti = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        ti = ti == 0 ? 0 : std::min(ti - 1, tris.size() - 1);
        const Tri& T = tris[ti];
        double a = r.ud(), b = r.ud();
        double sa = std::sqrt(a);
        V3 x = (1.0 - sa) * T.v0 + (sa * (1.0 - b)) * T.v1 + (sa * b) * T.v2;
        V3 n = T.n;
        if (T.two_sided && r.ud() < 0.5) n = -1.0 * n;
        // incident direction: cosine-distributed about -n, pointing into the surface
        V3 t, bt;
        onb(n, t, bt);
        double u1 = std::min(r.ud(), 0.999), u2 = r.ud();
        double rr = std::sqrt(u1), phi = 2.0 * M_PI * u2;
        V3 h = (rr * std::cos(phi)) * t + (rr * std::sin(phi)) * bt + std::sqrt(1.0 - u1) * n;
        V3 w = unit(-1.0 * h);
        float L[3];
        for (int k = 0; k < 3; ++k) {
            double v = r.ud();
            L[k] = d->radiance_log_uniform ? (float)std::exp(llo + (lhi - llo) * v)
                                           : (float)(d->radiance_lo + (d->radiance_hi - d->radiance_lo) * v);
        }
        const bool anti = r.ud() < d->antiradiance_prob;
        int span = d->max_bounce - d->min_bounce + 1;
        int bnc = d->min_bounce + std::min(span - 1, (int)(r.ud() * span));
        pos3[3 * i + 0] = (float)x.x;
        pos3[3 * i + 1] = (float)x.y;
        pos3[3 * i + 2] = (float)x.z;
        nrm3[3 * i + 0] = (float)n.x;
        nrm3[3 * i + 1] = (float)n.y;
        nrm3[3 * i + 2] = (float)n.z;
        dir3[3 * i + 0] = (float)w.x;
        dir3[3 * i + 1] = (float)w.y;
        dir3[3 * i + 2] = (float)w.z;
        for (int k = 0; k < 3; ++k) {
            rad3[3 * i + k] = L[k];
            antirad3[3 * i + k] = anti ? L[k] : 0.f;
        }
        material[i] = T.material;
        bounce[i] = bnc;
    }
    return 0;
}

int arc_scene_points(uint64_t seed, int n, int num_gaussians, float sigma, float* pos3) {
    if (n < 0 || !pos3 || (n - n / 2 > 0 && num_gaussians <= 0)) return -1;
    const int half = n / 2;
    std::vector<V3> centres((size_t)std::max(num_gaussians, 0));
    for (int k = 0; k < num_gaussians; ++k) {
        Rng r(seed, S_CENTRES, (uint64_t)k);
        centres[k] = {r.ud(), r.ud(), r.ud()};
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, S_POINTS, (uint64_t)i);
        V3 p;
        if (i < half) {
            p = {r.ud(), r.ud(), r.ud()};
        } else {
            int k = std::min(num_gaussians - 1, (int)(r.ud() * num_gaussians));
            double g[4];
            for (int j = 0; j < 2; ++j) {  // Box-Muller, two normals per draw
                double u1 = std::max(r.ud(), 1e-300), u2 = r.ud();
                double rad = std::sqrt(-2.0 * std::log(u1));
                g[2 * j] = rad * std::cos(2.0 * M_PI * u2);
                g[2 * j + 1] = rad * std::sin(2.0 * M_PI * u2);
            }
            p = centres[k] + (double)sigma * V3{g[0], g[1], g[2]};
        }
        pos3[3 * i + 0] = (float)p.x;
        pos3[3 * i + 1] = (float)p.y;
        pos3[3 * i + 2] = (float)p.z;
    }
    return 0;
}

int arc_scene_normals(uint64_t seed, int n, float* nrm3) {
    if (n < 0 || !nrm3) return -1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, S_NORMALS, (uint64_t)i);
        V3 v = uniform_sphere(r);
        nrm3[3 * i + 0] = (float)v.x;
        nrm3[3 * i + 1] = (float)v.y;
        nrm3[3 * i + 2] = (float)v.z;
    }
    return 0;
}

int arc_scene_num_triangles(const arc_scene_desc* d) {
    if (!valid(d)) return -1;
    int n = 0;
    for (const Tri& t : build_scene(*d)) n += t.two_sided ? 2 : 1;
    return n;
}

int arc_scene_triangles(const arc_scene_desc* d, float* p0, float* p1, float* p2, int* material, int max_tris) {
    if (!valid(d) || !p0 || !p1 || !p2 || !material) return -1;
    const std::vector<Tri> tris = build_scene(*d);
    int k = 0;
    auto put = [&](V3 a, V3 b, V3 c, int mat) {
        if (k >= max_tris) return;
        const V3 v[3] = {a, b, c};
        float* o[3] = {p0, p1, p2};
        for (int j = 0; j < 3; ++j) {
            o[j][3 * k + 0] = (float)v[j].x;
            o[j][3 * k + 1] = (float)v[j].y;
            o[j][3 * k + 2] = (float)v[j].z;
        }
        material[k] = mat;
        ++k;
    };
    for (const Tri& t : tris) {
        // the front face's normal is normalize(cross(p2 - p0, p1 - p0))
        const bool keep = dot(cross(t.v2 - t.v0, t.v1 - t.v0), t.n) > 0.0;
        if (keep) put(t.v0, t.v1, t.v2, t.material);
        else put(t.v0, t.v2, t.v1, t.material);
        if (t.two_sided) {  // the back face as a triangle of its own
            if (keep) put(t.v0, t.v2, t.v1, t.material);
            else put(t.v0, t.v1, t.v2, t.material);
        }
    }
    return k;
}

int arc_scene_light(const arc_scene_desc* d, float* center3, float* front3, float* up3, float* size2,
                    float* radiance3, float* material10) {
    if (!valid(d)) return -1;
    // a Cornell-style ceiling light just below the ceiling, facing down
    const float c[3] = {0.5f, 0.999f, 0.5f}, f[3] = {0.f, -1.f, 0.f}, u[3] = {0.f, 0.f, 1.f};
    const float rad[3] = {17.f, 12.f, 4.f};
    for (int k = 0; k < 3; ++k) {
        if (center3) center3[k] = c[k];
        if (front3) front3[k] = f[k];
        if (up3) up3[k] = u[k];
        if (radiance3) radiance3[k] = rad[k];
    }
    if (size2) {
        size2[0] = 130.f / 556.f;
        size2[1] = 105.f / 559.2f;
    }
    if (material10) {  // MATERIAL with emissive = radiance (AreaLight.h, "LightSource")
        for (int k = 0; k < 10; ++k) material10[k] = 0.f;
        for (int k = 0; k < 3; ++k) material10[k] = rad[k];
    }
    return 0;
}

}  // extern "C"
