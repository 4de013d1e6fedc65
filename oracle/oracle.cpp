// CPU restatement of the AntiradianceCuts hot path — TEST INFRASTRUCTURE ONLY.
//
// Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
// as the checker or the timed CPU baseline; never by the product path.
// Build: g++ -O2 -ffp-contract=off -fopenmp (oracle/Makefile).  No fast-math:
// REF32 needs IEEE-754 division and square root in the reference's order.

#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include <omp.h>

namespace {

// ===================================================================== LBVH

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// delta (bvh.cu:124-135): -1 outside [0, num_nodes]; common prefix of the
// codes, extended by the common prefix of the slot indices on equal codes.
inline int delta(int i, int j, const uint32_t* m, int num_nodes) {
    if (j < 0 || j > num_nodes) return -1;
    int d = clz32(m[i] ^ m[j]);
    if (d == 32) d += clz32((uint32_t)(i ^ j));
    return d;
}

inline int sgn(int v) { return v < 0 ? -1 : 1; }  // bvh.cu:170-173, sign(0) = +1

// One internal node of the Karras tree (bvh.cu:188-266).
void radix_node(int i, const uint32_t* m, int num_nodes, orc_node* nodes, int* parents, int* ranges) {
    const int root_slot = num_nodes + 1;  // parents[numLeafs + k] holds internal node k
    const int d = sgn(delta(i, i + 1, m, num_nodes) - delta(i, i - 1, m, num_nodes));
    const int dmin = delta(i, i - d, m, num_nodes);
    // upper bound for the range length: starts at 128, grows by 4x (bvh.cu:208-211)
    int lmax = 128;
    while (delta(i, i + lmax * d, m, num_nodes) > dmin) lmax *= 4;
    // binary search for the far end (bvh.cu:214-222)
    int l = 0;
    for (int t = lmax / 2; t > 0; t /= 2)
        if (delta(i, i + (l + t) * d, m, num_nodes) > dmin) l += t;
    const int j = i + l * d;
    // split search with t = ceil(l/2), ceil(l/4), ..., 1 (bvh.cu:225-239)
    const int dnode = delta(i, j, m, num_nodes);
    int s = 0;
    for (int k = 2;; k *= 2) {
        const int t = (l + k - 1) / k;
        if (delta(i, i + (s + t) * d, m, num_nodes) > dnode) s += t;
        if (t <= 1) break;
    }
    const int split = i + s * d + std::min(d, 0);
    const int lo = std::min(i, j), hi = std::max(i, j);
    orc_node nd;
    std::memset(&nd, 0, sizeof(nd));
    nd.left = (lo == split) ? split : -(split + 1);
    nd.right = (hi == split + 1) ? (split + 1) : -(split + 2);
    nd.visited = 0;
    parents[nd.left >= 0 ? nd.left : root_slot - (nd.left + 1)] = i;
    parents[nd.right >= 0 ? nd.right : root_slot - (nd.right + 1)] = i;
    nodes[i].left = nd.left;
    nodes[i].right = nd.right;
    nodes[i].visited = 0;
    if (ranges) {
        ranges[2 * i] = lo;
        ranges[2 * i + 1] = hi;
    }
    if (i == 0) parents[root_slot] = -1;
}

struct Rep {
    // AvplBvhNodeDataParam (bvh.h:104-113) over 2n-1 slots
    int* size;
    int* mat;
    const float* rnd;
    float* pos;
    float* nrm;
    float* rad;
    float* dir;
};

inline float len3f(const float* v) {
    // length() of cudaUtil.hpp:1420-1423: sqrtf(dot(v, v)), dot left to right
    return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// cluster() of bvh.cu:726-744.
void cluster(int target, int left, int right, int n, Rep& r) {
    const int tid = n + target;
    const int lid = left < 0 ? n - (left + 1) : left;
    const int rid = right < 0 ? n - (right + 1) : right;
    const float i1 = len3f(r.rad + 3 * (size_t)lid);
    const float i2 = len3f(r.rad + 3 * (size_t)rid);
    int repr = lid;
    if (r.rnd[tid] > (i1 / (i1 + i2))) repr = rid;  // NaN ratio keeps the left child (Q4)
    r.size[tid] = r.size[lid] + r.size[rid];
    r.mat[tid] = r.mat[repr];
    for (int k = 0; k < 3; ++k) {
        r.rad[3 * (size_t)tid + k] = r.rad[3 * (size_t)lid + k] + r.rad[3 * (size_t)rid + k];
        r.dir[3 * (size_t)tid + k] = r.dir[3 * (size_t)repr + k];
        r.pos[3 * (size_t)tid + k] = r.pos[3 * (size_t)repr + k];
        r.nrm[3 * (size_t)tid + k] = r.nrm[3 * (size_t)repr + k];
    }
}

// kernel_innerNodes (bvh.cu:280-316) run leaf by leaf in slot order: the
// first arrival at a node stops, the second merges both children.
void refit(orc_node* nodes, const int* parents, const float* sp, int n, Rep* rep) {
    for (int i = 0; i < n - 1; ++i) nodes[i].visited = 0;
    for (int i = 0; i < n; ++i) {
        int parent = parents[i];
        while (true) {
            if (nodes[parent].visited++ == 0) break;
            const int l = nodes[parent].left, r = nodes[parent].right;
            const float* lmin = l < 0 ? nodes[-(l + 1)].bbMin : sp + 3 * (size_t)l;
            const float* lmax = l < 0 ? nodes[-(l + 1)].bbMax : sp + 3 * (size_t)l;
            const float* rmin = r < 0 ? nodes[-(r + 1)].bbMin : sp + 3 * (size_t)r;
            const float* rmax = r < 0 ? nodes[-(r + 1)].bbMax : sp + 3 * (size_t)r;
            for (int k = 0; k < 3; ++k) {
                nodes[parent].bbMin[k] = std::fmin(lmin[k], rmin[k]);
                nodes[parent].bbMax[k] = std::fmax(lmax[k], rmax[k]);
            }
            if (rep) cluster(parent, l, r, n, *rep);
            parent = parents[n + parent];
            if (parent == -1) break;
        }
    }
}

// Bvh::create steps shared by SimpleBvh and AvplBvh (bvh.cu:325-374).
void codes_and_order(const float* pos3, const float* nrm3, int n, int consider_normals,
                     std::vector<uint32_t>& code, std::vector<int>& order) {
    float pmn[3], pmx[3], nmn[3], nmx[3];
    orc_bounds(pos3, n, pmn, pmx);
    std::vector<float> np((size_t)3 * n), nn((size_t)3 * n);
    orc_normalize(pos3, n, pmn, pmx, np.data());
    if (nrm3) {
        orc_bounds(nrm3, n, nmn, nmx);
        orc_normalize(nrm3, n, nmn, nmx, nn.data());
    }
    code.resize(n);
    for (int i = 0; i < n; ++i) {
        const float* p = &np[3 * (size_t)i];
        if (consider_normals) {
            const float* q = &nn[3 * (size_t)i];
            // MC_SIZE is uint32_t: the 60-bit code is truncated (bvh.h:17, bvh.cu:364-365)
            code[i] = (uint32_t)orc_morton6d(p[0], p[1], p[2], q[0], q[1], q[2]);
        } else {
            code[i] = orc_morton3d(p[0], p[1], p[2]);
        }
    }
    // thrust::sort_by_key (bvh.cu:795-807, 888-894): stable order on the code
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });
}

// ================================================================== gather

template <typename T>
struct V {
    T x, y, z;
};
template <typename T> inline V<T> mk(T x, T y, T z) { return {x, y, z}; }
template <typename T> inline V<T> ld(const float* p) { return {(T)p[0], (T)p[1], (T)p[2]}; }
template <typename T> inline V<T> operator+(V<T> a, V<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V<T> operator-(V<T> a, V<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V<T> operator-(V<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V<T> operator*(V<T> a, V<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> inline V<T> operator*(V<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V<T> operator*(T s, V<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V<T> operator/(V<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> inline T dot(V<T> a, V<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline T length(V<T> a) { return std::sqrt(dot(a, a)); }
// normalize of cudaUtil.hpp:1438-1442 (v * rsqrtf(dot(v,v))), rsqrt as 1/sqrt
template <typename T> inline V<T> normalize(V<T> a) { return a * (T(1) / std::sqrt(dot(a, a))); }
template <typename T> inline T clampv(T f, T a, T b) { return std::fmax(a, std::fmin(f, b)); }
// reflect of cudaUtil.hpp:1540-1543
template <typename T> inline V<T> reflect(V<T> i, V<T> n) { return i - T(2) * n * dot(n, i); }

// cudaUtil.hpp:46-48 constants (float literals, used in both modes)
const float kPi = 3.14159265f;
const float kOneOverPi = 0.318309886f;

template <typename T>
struct Mat {
    V<T> diffuse, specular;
    T exponent;
};

template <typename T>
Mat<T> material(const orc_scene* s, int idx) {
    Mat<T> m;
    const float* e = s->mats + 10 * (size_t)idx;
    m.diffuse = ld<T>(e + 3);
    m.specular = ld<T>(e + 6);
    m.exponent = (T)e[9];
    return m;
}

// G (cudaGather.cu:45-57)
template <typename T>
T Gterm(V<T> p1, V<T> n1, V<T> p2, V<T> n2) {
    V<T> a = normalize(n1), b = normalize(n2), w = normalize(p2 - p1);
    T c1 = clampv(dot(a, w), T(0), T(1));
    T c2 = clampv(dot(b, -w), T(0), T(1));
    T dist = length(p2 - p1);
    return (c1 * c2) / (dist * dist);
}

// f_r (cudaGather.cu:59-68)
template <typename T>
V<T> f_r(V<T> wi, V<T> wo, V<T> n, const Mat<T>& m) {
    V<T> d = T(kOneOverPi) * m.diffuse;
    V<T> s = mk<T>(0, 0, 0);
    if (length(m.specular) > T(0)) {
        T c = std::fmax(T(0), dot(reflect(-wi, n), wo));
        s = (T(0.5f) * T(kOneOverPi) * (m.exponent + T(2)) * std::pow(c, m.exponent)) * m.specular;
    }
    return d + s;
}

// f_r(from, over, to) (cudaGather.cu:70-76)
template <typename T>
V<T> f_r3(V<T> from, V<T> over, V<T> to, V<T> n, const Mat<T>& m) {
    return f_r(normalize(from - over), normalize(to - over), n, m);
}

// getRadiance (cudaGather.cu:78-95)
template <typename T>
V<T> radiance(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> anrm, V<T> aw, V<T> aL,
              const Mat<T>& amat, V<T> cam) {
    V<T> dir = normalize(pos - apos);
    V<T> brdf_light = f_r(-aw, dir, anrm, amat);
    if (length(aw) == T(0)) brdf_light = mk<T>(1, 1, 1);  // light-source AVPL
    V<T> brdf = f_r3(apos, pos, cam, nrm, mat);
    return aL * brdf_light * Gterm(pos, nrm, apos, anrm) * brdf;
}

// getAntiradiance, disk model (cudaGather.cu:97-144)
template <typename T>
V<T> antirad_disk(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> anrm, V<T> aw, V<T> aA,
                  V<T> cam, T radius) {
    V<T> zero = mk<T>(0, 0, 0);
    const T denom = dot(aw, nrm);
    if (denom == T(0)) return zero;
    const T t = dot(pos - apos, nrm) / denom;
    if (t <= T(0)) return zero;
    V<T> hit = apos + t * normalize(aw);
    if (length(hit - pos) >= radius) return zero;
    if (dot(normalize(hit - apos), aw) <= T(0)) return zero;
    if (dot(nrm, aw) >= T(0)) return zero;
    V<T> dir = normalize(pos - apos);
    if (dot(dir, anrm) >= T(0)) return zero;
    V<T> brdf = f_r3(apos, pos, cam, nrm, mat);
    return aA * brdf / (T(kPi) * radius * radius);
}

// GetAntiradiance, cone model (Shaders/Gather.frag:149-188 with filter mode 0
// and no clamp) times the pixel BRDF, as in the call at Gather.frag:134.
template <typename T>
V<T> antirad_cone(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> aw, V<T> aA, V<T> cam,
                  T half_angle) {
    V<T> zero = mk<T>(0, 0, 0);
    const T eps = T(0.001f);
    V<T> to_pos = normalize(pos - apos);
    T dist = length(pos - apos);
    V<T> wn = normalize(aw);
    T c = dot(to_pos, wn);
    if (c <= eps) return zero;
    T cp = dot(nrm, -to_pos);
    if (cp <= eps) return zero;
    if (!(std::acos(std::fmin(c, T(1))) < half_angle)) return zero;
    T GA = cp / (dist * dist);
    V<T> brdf = f_r3(apos, pos, cam, nrm, mat);
    return brdf * (GA * aA);
}

template <typename T>
struct Pixel {
    V<T> pos, nrm;
    Mat<T> mat;
};

template <typename T>
Pixel<T> pixel_of(const orc_scene* s, int p) {
    Pixel<T> px;
    px.pos = ld<T>(s->pos4 + 4 * (size_t)p);
    px.nrm = ld<T>(s->nrm4 + 4 * (size_t)p);
    int mi = (int)s->nrm4[4 * (size_t)p + 3];
    px.mat = material<T>(s, mi);
    if (s->albedo4) px.mat.diffuse = ld<T>(s->albedo4 + 4 * (size_t)p);
    return px;
}

template <typename T>
void pair(const orc_scene* s, const Pixel<T>& px, int i, V<T>& r, V<T>& a) {
    const V<T> cam = mk<T>(s->cam[0], s->cam[1], s->cam[2]);
    const V<T> apos = ld<T>(s->apos3 + 3 * (size_t)i), anrm = ld<T>(s->anrm3 + 3 * (size_t)i);
    const V<T> aw = ld<T>(s->adir3 + 3 * (size_t)i);
    const Mat<T> am = material<T>(s, s->amat[i]);
    r = radiance(px.pos, px.nrm, px.mat, apos, anrm, aw, ld<T>(s->arad3 + 3 * (size_t)i), am, cam);
    const V<T> A = ld<T>(s->aanti3 + 3 * (size_t)i);
    if (s->antirad_model == ORC_ANTIRAD_CONE)
        a = antirad_cone(px.pos, px.nrm, px.mat, apos, aw, A, cam, (T)s->cone_half_angle);
    else
        a = antirad_disk(px.pos, px.nrm, px.mat, apos, anrm, aw, A, cam, (T)s->photon_radius);
}

template <typename T>
void gather_exact(const orc_scene* s, const int* pixels, int npix, double* rad3, double* anti3, int what) {
    const bool do_rad = what & 1, do_anti = what & 2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < npix; ++q) {
        const int p = pixels ? pixels[q] : q;
        const Pixel<T> px = pixel_of<T>(s, p);
        const V<T> cam = mk<T>(s->cam[0], s->cam[1], s->cam[2]);
        V<T> R = mk<T>(0, 0, 0), A = mk<T>(0, 0, 0);
        for (int i = 0; i < s->num_avpls; ++i) {
            const V<T> apos = ld<T>(s->apos3 + 3 * (size_t)i), anrm = ld<T>(s->anrm3 + 3 * (size_t)i);
            const V<T> aw = ld<T>(s->adir3 + 3 * (size_t)i);
            if (do_rad) {
                const Mat<T> am = material<T>(s, s->amat[i]);
                R = R + radiance(px.pos, px.nrm, px.mat, apos, anrm, aw, ld<T>(s->arad3 + 3 * (size_t)i), am, cam);
            }
            if (do_anti) {
                const V<T> Av = ld<T>(s->aanti3 + 3 * (size_t)i);
                if (s->antirad_model == ORC_ANTIRAD_CONE)
                    A = A + antirad_cone(px.pos, px.nrm, px.mat, apos, aw, Av, cam, (T)s->cone_half_angle);
                else
                    A = A + antirad_disk(px.pos, px.nrm, px.mat, apos, anrm, aw, Av, cam, (T)s->photon_radius);
            }
        }
        rad3[3 * (size_t)q + 0] = R.x;
        rad3[3 * (size_t)q + 1] = R.y;
        rad3[3 * (size_t)q + 2] = R.z;
        anti3[3 * (size_t)q + 0] = A.x;
        anti3[3 * (size_t)q + 1] = A.y;
        anti3[3 * (size_t)q + 2] = A.z;
    }
}

// kernel_bvh (cudaGather.cu:316-444) for one pixel.
template <typename T>
void cut_pixel(const orc_scene* s, int p, int n, const orc_node* nodes, const int* nd_mat,
               const float* nd_pos3, const float* nd_nrm3, const float* nd_rad3, const float* nd_dir3,
               float ref_thresh, double* out3, int* err, long long& visited, long long& evaluated) {
    const Pixel<T> px = pixel_of<T>(s, p);
    const V<T> cam = mk<T>(s->cam[0], s->cam[1], s->cam[2]);
    auto eval = [&](int idx) {
        const Mat<T> am = material<T>(s, nd_mat[idx]);
        return radiance(px.pos, px.nrm, px.mat, ld<T>(nd_pos3 + 3 * (size_t)idx),
                        ld<T>(nd_nrm3 + 3 * (size_t)idx), ld<T>(nd_dir3 + 3 * (size_t)idx),
                        ld<T>(nd_rad3 + 3 * (size_t)idx), am, cam);
    };
    const int kStack = 32;  // STACK_SIZE (cudaGather.cu:19)
    int stack[kStack];
    stack[0] = -1;
    int sp = 1;
    bool error = false;
    V<T> R = mk<T>(0, 0, 0);
    while (sp > 0) {
        const int node = stack[--sp];
        if (node >= 0) {  // leaf
            R = R + eval(node);
            evaluated++;
            continue;
        }
        const int k = -(node + 1);
        if (k < 0 || k >= n - 1) {
            error = true;
            break;
        }
        const orc_node& b = nodes[k];
        const int idx = n + k;
        visited++;
        const V<T> bmax = ld<T>(b.bbMax), bmin = ld<T>(b.bbMin);
        const bool vis = dot(px.nrm, normalize(bmax - px.pos)) > T(0) ||
                         dot(px.nrm, normalize(bmin - px.pos)) > T(0);
        if (!vis) continue;
        const V<T> rep = ld<T>(nd_pos3 + 3 * (size_t)idx);
        const T radius = T(0.5f) * length(bmax - bmin);
        const T dist = length(rep - px.pos);
        const T omega = (T(2.f) * T(kPi)) * (T(1.f) - dist / std::sqrt(radius * radius + dist * dist));
        if (omega < (T)ref_thresh) {
            R = R + eval(idx);
            evaluated++;
        } else {
            stack[sp++] = b.left;
            stack[sp++] = b.right;
            if (sp >= kStack) {
                error = true;
                break;
            }
        }
    }
    if (error) {  // sentinel (1e5, 0, 1e5) of cudaGather.cu:440-442
        out3[0] = 100000.0;
        out3[1] = 0.0;
        out3[2] = 100000.0;
    } else {
        out3[0] = R.x;
        out3[1] = R.y;
        out3[2] = R.z;
    }
    *err = error ? 1 : 0;
}

}  // namespace

extern "C" {

// ================================================================== morton.h
uint32_t orc_expand_bits_by1(uint32_t x) {
    x &= 0x0000ffffu;
    x = (x ^ (x << 8)) & 0x00ff00ffu;
    x = (x ^ (x << 4)) & 0x0f0f0f0fu;
    x = (x ^ (x << 2)) & 0x33333333u;
    x = (x ^ (x << 1)) & 0x55555555u;
    return x;
}

uint32_t orc_expand_bits_by2(uint32_t x) {
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0xc30c30c3u;
    x = (x ^ (x << 2)) & 0x49249249u;
    return x;
}

uint32_t orc_interleave3(uint32_t x, uint32_t y, uint32_t z) {
    return (orc_expand_bits_by2(z) << 2) + (orc_expand_bits_by2(y) << 1) + orc_expand_bits_by2(x);
}

uint64_t orc_interleave6(uint64_t x, uint64_t y, uint64_t z, uint64_t u, uint64_t v, uint64_t w) {
    // bit i of each input lands at 6*i + {0..5} (morton.h:67-80)
    uint64_t r = 0;
    const uint64_t one = 1;
    for (int i = 0, k = 0; i < 10; ++i, k += 5) {
        r |= (x & one << i) << k | (y & one << i) << (k + 1) | (z & one << i) << (k + 2) |
             (u & one << i) << (k + 3) | (v & one << i) << (k + 4) | (w & one << i) << (k + 5);
    }
    return r;
}

static inline float quant10(float c) {
    // fminf(fmaxf(c * 1024, 0), 1023): NaN maps to 0 through fmaxf
    return std::fmin(std::fmax(c * 1024.0f, 0.0f), 1023.0f);
}

uint32_t orc_morton3d(float x, float y, float z) {
    return orc_interleave3((uint32_t)quant10(x), (uint32_t)quant10(y), (uint32_t)quant10(z));
}

uint64_t orc_morton6d(float ax, float ay, float az, float bx, float by, float bz) {
    return orc_interleave6((uint64_t)quant10(ax), (uint64_t)quant10(ay), (uint64_t)quant10(az),
                           (uint64_t)quant10(bx), (uint64_t)quant10(by), (uint64_t)quant10(bz));
}

// ==================================================================== bvh.cu
void orc_bounds(const float* p3, int n, float* mn3, float* mx3) {
    for (int k = 0; k < 3; ++k) {
        mn3[k] = std::numeric_limits<float>::max();
        mx3[k] = std::numeric_limits<float>::min();  // FLT_MIN, not -FLT_MAX (Q1)
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            mn3[k] = std::fmin(mn3[k], p3[3 * (size_t)i + k]);
            mx3[k] = std::fmax(mx3[k], p3[3 * (size_t)i + k]);
        }
}

void orc_normalize(const float* p3, int n, const float* mn3, const float* mx3, float* out3) {
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            out3[3 * (size_t)i + k] = (p3[3 * (size_t)i + k] - mn3[k]) / (mx3[k] - mn3[k]);
}

int orc_delta(int i, int j, const uint32_t* codes, int num_nodes) { return delta(i, j, codes, num_nodes); }

int orc_radix_tree(const uint32_t* codes, int n, orc_node* nodes, int* parents, int* ranges) {
    if (n <= 1) return -1;
    for (int i = 0; i < n - 1; ++i) {
        std::memset(&nodes[i], 0, sizeof(orc_node));
        radix_node(i, codes, n - 1, nodes, parents, ranges);
    }
    return 0;
}

int orc_refit(orc_node* nodes, const int* parents, const float* sorted_pos3, int n) {
    if (n <= 1) return -1;
    refit(nodes, parents, sorted_pos3, n, nullptr);
    return 0;
}

int orc_check_tree_integrity(const orc_node* nodes, const int* parents, int n, int* max_stack) {
    if (n <= 1) return -1;
    const int num_nodes = n - 1, num_leafs = n;
    std::vector<int> stack;
    stack.push_back(-1);
    std::vector<int> checkparent(num_nodes, 0), visited(num_nodes + num_leafs, 0);
    int bad = 0;
    size_t maxsp = 0;
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        int slot;
        if (node >= 0) {
            if (node >= num_leafs) return 1;
            slot = node;
        } else {
            const int k = -(node + 1);
            if (k >= num_nodes) return 1;
            slot = num_leafs + k;
            stack.push_back(nodes[k].left);
            stack.push_back(nodes[k].right);
            maxsp = std::max(maxsp, stack.size());
        }
        const int parent = parents[slot];
        if (parent != -1) {
            if (parent < 0 || parent >= num_nodes) return 2;
            checkparent[parent]++;
        }
        if (visited[slot]++) return 3;  // visited twice
        if (stack.size() > (size_t)4 * (num_nodes + num_leafs)) return 4;
    }
    for (int c : checkparent) bad |= (c != 2);
    if (bad) return 5;
    for (int v : visited)
        if (v != 1) return 6;
    // parent pointers agree with child pointers
    for (int k = 0; k < num_nodes; ++k) {
        const int l = nodes[k].left, r = nodes[k].right;
        const int ls = l >= 0 ? l : num_leafs + (-(l + 1));
        const int rs = r >= 0 ? r : num_leafs + (-(r + 1));
        if (parents[ls] != k || parents[rs] != k) return 7;
    }
    if (parents[num_leafs] != -1) return 8;
    if (max_stack) *max_stack = (int)maxsp;
    return 0;
}

int orc_lbvh_build(const float* pos3, const float* nrm3, int n, int consider_normals,
                   uint32_t* codes, int* perm, orc_node* nodes, int* parents, int* ranges,
                   float* sorted_pos3, float* sorted_nrm3) {
    if (n <= 1) return -1;
    std::vector<uint32_t> code;
    std::vector<int> order;
    codes_and_order(pos3, nrm3, n, consider_normals, code, order);
    std::vector<float> sp((size_t)3 * n);
    for (int s = 0; s < n; ++s) {
        codes[s] = code[order[s]];
        if (perm) perm[s] = order[s];
        for (int k = 0; k < 3; ++k) {
            sp[3 * (size_t)s + k] = pos3[3 * (size_t)order[s] + k];
            if (sorted_nrm3 && nrm3) sorted_nrm3[3 * (size_t)s + k] = nrm3[3 * (size_t)order[s] + k];
        }
    }
    if (sorted_pos3) std::memcpy(sorted_pos3, sp.data(), sp.size() * sizeof(float));
    orc_radix_tree(codes, n, nodes, parents, ranges);
    refit(nodes, parents, sp.data(), n, nullptr);
    return 0;
}

int orc_avpl_bvh_build(int n, const float* pos3, const float* nrm3, const float* dir3,
                       const float* rad3, const int* mat, const int* bounce, const float* rand,
                       uint32_t* codes, int* perm, orc_node* nodes, int* parents, int* ranges,
                       float* sorted_pos3, float* sorted_nrm3,
                       int* nd_size, int* nd_mat, float* nd_pos3, float* nd_nrm3,
                       float* nd_rad3, float* nd_dir3) {
    if (n <= 1) return -1;
    for (int i = 0; i < n; ++i)
        if (bounce[i] <= 0) return -2;  // leaves and node data would misalign (Q2)
    std::vector<uint32_t> code;
    std::vector<int> order;
    // AvplBvh passes considerNormals = false from CudaGather (cudaGather.cu:858)
    codes_and_order(pos3, nrm3, n, 0, code, order);
    std::vector<float> sp((size_t)3 * n);
    for (int s = 0; s < n; ++s) {
        const int o = order[s];
        codes[s] = code[o];
        if (perm) perm[s] = o;
        nd_size[s] = 1;
        nd_mat[s] = mat[o];
        for (int k = 0; k < 3; ++k) {
            sp[3 * (size_t)s + k] = pos3[3 * (size_t)o + k];
            if (sorted_nrm3) sorted_nrm3[3 * (size_t)s + k] = nrm3[3 * (size_t)o + k];
            nd_pos3[3 * (size_t)s + k] = pos3[3 * (size_t)o + k];
            nd_nrm3[3 * (size_t)s + k] = nrm3[3 * (size_t)o + k];
            nd_rad3[3 * (size_t)s + k] = rad3[3 * (size_t)o + k];
            nd_dir3[3 * (size_t)s + k] = dir3[3 * (size_t)o + k];
        }
    }
    if (sorted_pos3) std::memcpy(sorted_pos3, sp.data(), sp.size() * sizeof(float));
    orc_radix_tree(codes, n, nodes, parents, ranges);
    Rep rep{nd_size, nd_mat, rand, nd_pos3, nd_nrm3, nd_rad3, nd_dir3};
    refit(nodes, parents, sp.data(), n, &rep);
    return 0;
}

// ============================================================ cudaGather.cu
int orc_pair(const orc_scene* s, int precision, int pixel, int avpl, double* rad3, double* anti3) {
    if (precision == 64) {
        Pixel<double> px = pixel_of<double>(s, pixel);
        V<double> r, a;
        pair<double>(s, px, avpl, r, a);
        rad3[0] = r.x; rad3[1] = r.y; rad3[2] = r.z;
        anti3[0] = a.x; anti3[1] = a.y; anti3[2] = a.z;
    } else {
        Pixel<float> px = pixel_of<float>(s, pixel);
        V<float> r, a;
        pair<float>(s, px, avpl, r, a);
        rad3[0] = r.x; rad3[1] = r.y; rad3[2] = r.z;
        anti3[0] = a.x; anti3[1] = a.y; anti3[2] = a.z;
    }
    return 0;
}

int orc_gather_exact(const orc_scene* s, int precision, const int* pixels, int npix,
                     double* rad3, double* anti3) {
    return orc_gather_exact_parts(s, precision, pixels, npix, rad3, anti3, 3);
}

int orc_gather_exact_parts(const orc_scene* s, int precision, const int* pixels, int npix,
                           double* rad3, double* anti3, int what) {
    if (!s || npix < 0) return -1;
    if (precision == 64)
        gather_exact<double>(s, pixels, npix, rad3, anti3, what);
    else
        gather_exact<float>(s, pixels, npix, rad3, anti3, what);
    return 0;
}

int orc_gather_cut_solid_angle(const orc_scene* s, int precision, int n_leaves,
                               const orc_node* nodes, const int* nd_mat, const float* nd_pos3,
                               const float* nd_nrm3, const float* nd_rad3, const float* nd_dir3,
                               float ref_thresh, const int* pixels, int npix,
                               double* rad3, int* err, long long* visited, long long* evaluated) {
    if (!s || n_leaves <= 1 || npix < 0) return -1;
    long long vis = 0, ev = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : vis, ev)
    for (int q = 0; q < npix; ++q) {
        const int p = pixels ? pixels[q] : q;
        if (precision == 64)
            cut_pixel<double>(s, p, n_leaves, nodes, nd_mat, nd_pos3, nd_nrm3, nd_rad3, nd_dir3,
                              ref_thresh, rad3 + 3 * (size_t)q, err + q, vis, ev);
        else
            cut_pixel<float>(s, p, n_leaves, nodes, nd_mat, nd_pos3, nd_nrm3, nd_rad3, nd_dir3,
                             ref_thresh, rad3 + 3 * (size_t)q, err + q, vis, ev);
    }
    if (visited) *visited = vis;
    if (evaluated) *evaluated = ev;
    return 0;
}

}  // extern "C"

// ---- output side --------------------------------------------------------
extern "C" {

void orc_image_accumulate(float* accum4, const float* frame4, int npix) {  // Render.cpp:788-806
    for (int i = 0; i < npix; ++i) {
        for (int c = 0; c < 3; ++c) accum4[4 * i + c] = accum4[4 * i + c] + frame4[4 * i + c];
        accum4[4 * i + 3] = accum4[4 * i + 3] + 1.f;
    }
}

void orc_image_normalize(float* out4, const float* accum4, int npix, int num_paths) {  // Render.cpp:642-667
    const float inv = num_paths != 0 ? 1.f / (float)num_paths : 0.f;
    for (int i = 0; i < npix; ++i) {
        for (int c = 0; c < 3; ++c) out4[4 * i + c] = accum4[4 * i + c] * inv;
        out4[4 * i + 3] = 1.f;
    }
}

float orc_image_error(const float* ref4, const float* img4, int npix, float* e3) {  // CReferenceImage.cpp:28-51
    float e[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < npix; ++i)
        for (int c = 0; c < 3; ++c) {
            const float d = ref4[4 * i + c] - img4[4 * i + c];
            e[c] += std::min(100.f, d * d);
        }
    const float inv = 1.f / (float)npix;
    for (int c = 0; c < 3; ++c) e[c] = std::sqrt(inv * e[c]);
    if (e3)
        for (int c = 0; c < 3; ++c) e3[c] = e[c];
    return 1.f / 3.f * (e[0] * e[1] + e[2]);
}

}  // extern "C"

// ---- ray tracing and the AVPL shooter ----------------------------------
namespace {

struct T3 {
    float x, y, z;
};
inline T3 t3(const float* p) { return T3{p[0], p[1], p[2]}; }
inline T3 operator+(T3 a, T3 b) { return T3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline T3 operator-(T3 a, T3 b) { return T3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline T3 operator*(float s, T3 a) { return T3{s * a.x, s * a.y, s * a.z}; }
inline T3 mul(T3 a, T3 b) { return T3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot3(T3 a, T3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline T3 cross3(T3 a, T3 b) { return T3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float len3(T3 a) { return std::sqrt(dot3(a, a)); }
inline T3 norm3(T3 a) { return (1.0f / len3(a)) * a; }

bool tri_intersect(T3 P0, T3 P1, T3 P2, T3 o, T3 d, float det_eps, float& t) {  // Triangle.h, FRONT_FACE
    const T3 e1 = P2 - P0, e2 = P1 - P0;
    const T3 p = cross3(d, e2);
    const float det = dot3(e1, p);
    if (det < det_eps && det > -det_eps) return false;
    const T3 tt = o - P0;
    const float u = dot3(tt, p);
    if (u < 0.0f || u > det) return false;
    const T3 q = cross3(tt, e1);
    const float v = dot3(d, q);
    if (v < 0.0f || u + v > det) return false;
    t = 1.f / det * dot3(e2, q);
    return true;
}

// the library's counter-based uniform floats: (mix64 hash >> 40) * 2^-24
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
inline float hash_uniform(uint64_t seed, uint64_t i) {
    const uint64_t h = mix64(seed * 0x9e3779b97f4a7c15ULL + mix64(i + 0x632be59bd9b4e019ULL));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

const float kTwoPiF = 6.28318530718f, kOneOverTwoPiF = 0.15915494309f, kOneOverPiF = 0.318309886f;

T3 sample_cos_cone(T3 n, float u1, float u2, float order, float& pdf) {  // Sampler.h:13-35
    const T3 tt{n.y, n.z, -n.x};
    const T3 U = norm3(cross3(tt, n));
    const T3 W = norm3(cross3(U, n));
    // order 1: cos(theta) = sqrt(u1), sin(theta) = sqrt(1 - u1) (the value of
    // the reference's acos / pow / sin / cos up to rounding)
    float ct, st;
    if (order == 1.f) {
        ct = std::sqrt(u1);
        st = std::sqrt(std::max(0.f, 1.f - u1));
    } else {
        const float theta = std::acos(std::pow(u1, 1.0f / (order + 1.0f)));
        ct = std::cos(theta);
        st = std::sin(theta);
    }
    const float phi = kTwoPiF * u2;
    const T3 s = norm3(T3{st * std::cos(phi), st * std::sin(phi), ct});
    pdf = (order + 1.f) * kOneOverTwoPiF * std::pow(s.z, order);
    return norm3(T3{U.x * s.x + W.x * s.y + n.x * s.z, U.y * s.x + W.y * s.y + n.y * s.z,
                    U.z * s.x + W.z * s.y + n.z * s.z});
}

float lum(T3 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }

T3 phong_eval(T3 wi, T3 wo, T3 n, const float* m) {  // Brdf.h:9-22
    if (dot3(wi, n) <= 0.f || dot3(wo, n) <= 0.f) return T3{0, 0, 0};
    T3 res{0, 0, 0};
    const T3 spec = t3(m + 6);
    if (len3(spec) > 0.f) {
        const T3 r = norm3((2.f * dot3(wi, n)) * n - wi);
        const float c = std::max(0.f, dot3(r, wo));
        res = res + (kOneOverTwoPiF * (m[9] + 2.f) * std::pow(c, m[9])) * spec;
    }
    return res + kOneOverPiF * t3(m + 3);
}

T3 phong_sample(T3 wi, T3 n, const float* m, float s0, float s1, float& pdf) {  // Brdf.h:54-82
    const bool hd = len3(t3(m + 3)) > 0.f, hs = len3(t3(m + 6)) > 0.f;
    bool spec = hs;
    if (hd && hs) {
        const float a = lum(t3(m + 6)) / (lum(t3(m + 6)) + lum(t3(m + 3)));
        if (s0 <= a) {
            s0 /= a;
        } else {
            s0 = (s0 - a) / (1.f - a);
            spec = false;
        }
    }
    if (spec) {
        const T3 dir = sample_cos_cone(wi, s0, s1, m[9], pdf);
        if (dot3(dir, n) <= 0.f) return T3{0, 0, 0};
    }
    return sample_cos_cone(n, s0, s1, 1.f, pdf);
}

}  // namespace

extern "C" {

int orc_tri_intersect(const float* p0, const float* p1, const float* p2, const float* org, const float* dir,
                      float det_eps, float* t) {
    float tt = 0.f;
    const bool h = tri_intersect(t3(p0), t3(p1), t3(p2), t3(org), t3(dir), det_eps, tt);
    if (h && t) *t = tt;
    return h ? 1 : 0;
}

int orc_trace_closest(int ntri, const float* p0, const float* p1, const float* p2, const int* material,
                      const int* emissive, const float* org, const float* dir, float det_eps, float* t) {
    int best = -1;
    float tb = INFINITY;
    const T3 o = t3(org), d = t3(dir);
    for (int i = 0; i < ntri; ++i) {
        float tt;
        if (tri_intersect(t3(p0 + 3 * i), t3(p1 + 3 * i), t3(p2 + 3 * i), o, d, det_eps, tt) && tt > 0.f &&
            tt < tb) {
            tb = tt;
            best = i;
        }
    }
    if (best >= 0 && emissive && material && emissive[material[best] & 255]) best = -1;
    if (t) *t = best >= 0 ? tb : -1.f;
    return best;
}

int orc_trace_shoot(int ntri, const float* p0, const float* p1, const float* p2, const int* material,
                    const float* mats, int num_mats, const float* light, uint64_t seed, int first, int count,
                    int explicit_direct, int* counts, float* pos3, float* nrm3, float* dir3, float* rad3,
                    float* anti3, int* mat_out, int* bounce_out, int max_avpls) {
    std::vector<int> emissive(256, 0);
    for (int k = 0; k < num_mats && k < 256; ++k)
        emissive[k] = (mats[10 * k] * mats[10 * k] + mats[10 * k + 1] * mats[10 * k + 1] +
                       mats[10 * k + 2] * mats[10 * k + 2]) > 0.f;
    std::vector<float> zero(10, 0.f);
    auto material_of = [&](int m) { return (m & 255) < num_mats ? mats + 10 * (m & 255) : zero.data(); };
    const T3 lc = t3(light), lf = t3(light + 3), ls = t3(light + 6), lu = t3(light + 9);
    const float hw = light[12], hh = light[13];
    const T3 lrad = t3(light + 14);
    const int lmat = (int)light[17];
    const float area = (2.f * hw) * (2.f * hh);
    int total = 0;
    for (int path = first; path < first + count; ++path) {
        const uint64_t base = (uint64_t)path << 24;
        uint32_t k = 0;
        auto rnd = [&]() { return hash_uniform(seed, base | (uint64_t)(k++)); };
        const float sx = -1.f + 2.f * rnd();
        const float sy = -1.f + 2.f * rnd();
        T3 pos = lc + (sx * hw) * ls + (sy * hh) * lu;
        float pdf = 1.f / area;
        T3 nrm = lf;
        T3 rad = (1.f / pdf) * lrad;
        T3 in_dir{0, 0, 0};
        int bounce = 0, mat = lmat;
        const float u1 = rnd(), u2 = rnd();
        T3 direction = sample_cos_cone(nrm, u1, u2, 1.f, pdf);
        int n = 0;
        for (;;) {
            const T3 org = pos + 1e-4f * nrm;
            float t;
            const float o3[3] = {org.x, org.y, org.z}, d3[3] = {direction.x, direction.y, direction.z};
            const int s = orc_trace_closest(ntri, p0, p1, p2, material, emissive.data(), o3, d3, 1e-4f, &t);
            if (s < 0) break;
            const float cos_theta = dot3(nrm, direction);
            T3 contrib{rad.x * cos_theta / pdf, rad.y * cos_theta / pdf, rad.z * cos_theta / pdf};
            if (bounce > 0) contrib = mul(contrib, phong_eval(-1.f * in_dir, direction, nrm, material_of(mat)));
            const T3 P0 = t3(p0 + 3 * s), P1 = t3(p1 + 3 * s), P2 = t3(p2 + 3 * s);
            T3 hn = cross3(P2 - P0, P1 - P0);
            if (len3(hn) > 0.f) hn = norm3(hn);
            const T3 hit = org + t * direction;
            pos = hit - 1e-4f * hn;
            nrm = hn;
            rad = contrib;
            T3 anti = contrib;
            in_dir = direction;
            ++bounce;
            mat = material[s];
            if (bounce == 1 && explicit_direct) anti = T3{0, 0, 0};
            const bool stop = rnd() > 0.8f;
            rad = stop ? T3{0, 0, 0} : (1.f / 0.8f) * rad;
            if (pos3 && total < max_avpls) {
                const T3 vals[5] = {pos, nrm, in_dir, rad, anti};
                float* outs[5] = {pos3, nrm3, dir3, rad3, anti3};
                for (int a = 0; a < 5; ++a) {
                    outs[a][3 * total + 0] = vals[a].x;
                    outs[a][3 * total + 1] = vals[a].y;
                    outs[a][3 * total + 2] = vals[a].z;
                }
                mat_out[total] = mat;
                bounce_out[total] = bounce;
            }
            ++total;
            ++n;
            if (stop) break;
            const float s0 = rnd(), s1 = rnd();
            direction = phong_sample(-1.f * direction, nrm, material_of(mat), s0, s1, pdf);
            if (len3(direction) <= 0.f) break;
            if (pdf < 1e-4f) break;
        }
        if (counts) counts[path - first] = n;
    }
    return total;
}

}  // extern "C"

int orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
