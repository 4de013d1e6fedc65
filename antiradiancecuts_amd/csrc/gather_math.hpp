// Device-side pair functions of the AVPL gather (cudaGather.cu:34-154 and
// Shaders/Gather.frag:149-188), shared by the gather kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "arc_internal.hpp"

namespace arc {
namespace gm {

// cudaUtil.hpp:46-48 constants
constexpr float kPi = 3.14159265f;
constexpr float kOneOverPi = 0.318309886f;
constexpr float kConeEps = 0.001f;  // EPSILON of Shaders/Gather.frag:10

template <typename T>
struct V {
    T x, y, z;
};
template <typename T> __device__ inline V<T> mk(T x, T y, T z) { return {x, y, z}; }
template <typename T> __device__ inline V<T> ld3(const f3& a) { return {(T)a.x, (T)a.y, (T)a.z}; }
template <typename T> __device__ inline V<T> operator+(V<T> a, V<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> __device__ inline V<T> operator-(V<T> a, V<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ inline V<T> operator-(V<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> __device__ inline V<T> operator*(V<T> a, V<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> __device__ inline V<T> operator*(V<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> __device__ inline V<T> operator*(T s, V<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ inline V<T> operator/(V<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> __device__ inline T dot(V<T> a, V<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ inline float rsq(float x) { return rsqrtf(x); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }
template <typename T> __device__ inline T length(V<T> a) { return sqrt(dot(a, a)); }
template <typename T> __device__ inline V<T> normalize(V<T> a) { return a * rsq(dot(a, a)); }
template <typename T> __device__ inline T clamp01(T f) { return fmax(T(0), fmin(f, T(1))); }
template <typename T> __device__ inline V<T> reflect(V<T> i, V<T> n) { return i - T(2) * n * dot(n, i); }

template <typename T>
struct Mat {
    V<T> diffuse, specular;
    T exponent;
};

// The device table always holds kMaxMaterials entries (unused ones zero), so
// any index masked to that range stays inside it.
constexpr int kMaxMaterials = 256;
__device__ inline int mat_slot(int idx) { return idx & (kMaxMaterials - 1); }

template <typename T>
__device__ inline Mat<T> material(const arc_material* mats, int idx) {
    const arc_material& m = mats[mat_slot(idx)];
    Mat<T> r;
    r.diffuse = mk<T>(m.diffuse.x, m.diffuse.y, m.diffuse.z);
    r.specular = mk<T>(m.specular.x, m.specular.y, m.specular.z);
    r.exponent = (T)m.exponent;
    return r;
}

// G (cudaGather.cu:45-57)
template <typename T>
__device__ inline T Gterm(V<T> p1, V<T> n1, V<T> p2, V<T> n2) {
    const V<T> a = normalize(n1), b = normalize(n2), w = normalize(p2 - p1);
    const T c1 = clamp01(dot(a, w));
    const T c2 = clamp01(dot(b, -w));
    const T dist = length(p2 - p1);
    return (c1 * c2) / (dist * dist);
}

// f_r (cudaGather.cu:59-68)
template <typename T>
__device__ inline V<T> f_r(V<T> wi, V<T> wo, V<T> n, const Mat<T>& m) {
    V<T> d = T(kOneOverPi) * m.diffuse;
    if (length(m.specular) > T(0)) {
        const T c = fmax(T(0), dot(reflect(-wi, n), wo));
        d = d + (T(0.5f) * T(kOneOverPi) * (m.exponent + T(2)) * pow(c, m.exponent)) * m.specular;
    }
    return d;
}

template <typename T>
__device__ inline V<T> f_r3(V<T> from, V<T> over, V<T> to, V<T> n, const Mat<T>& m) {
    return f_r(normalize(from - over), normalize(to - over), n, m);
}

// getRadiance (cudaGather.cu:78-95)
template <typename T>
__device__ inline V<T> radiance(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> anrm, V<T> aw,
                                V<T> aL, const Mat<T>& amat, V<T> cam) {
    const V<T> dir = normalize(pos - apos);
    V<T> brdf_light = f_r(-aw, dir, anrm, amat);
    if (length(aw) == T(0)) brdf_light = mk<T>(1, 1, 1);
    const V<T> brdf = f_r3(apos, pos, cam, nrm, mat);
    return aL * brdf_light * Gterm(pos, nrm, apos, anrm) * brdf;
}

// getAntiradiance, disk model (cudaGather.cu:97-144)
template <typename T>
__device__ inline V<T> antirad_disk(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> anrm,
                                    V<T> aw, V<T> aA, V<T> cam, T radius) {
    const V<T> zero = mk<T>(0, 0, 0);
    const T denom = dot(aw, nrm);
    if (denom == T(0)) return zero;
    const T t = dot(pos - apos, nrm) / denom;
    if (t <= T(0)) return zero;
    const V<T> hit = apos + t * normalize(aw);
    if (length(hit - pos) >= radius) return zero;
    if (dot(normalize(hit - apos), aw) <= T(0)) return zero;
    if (dot(nrm, aw) >= T(0)) return zero;
    if (dot(normalize(pos - apos), anrm) >= T(0)) return zero;
    const V<T> brdf = f_r3(apos, pos, cam, nrm, mat);
    return aA * brdf / (T(kPi) * radius * radius);
}

// GetAntiradiance, cone model (Shaders/Gather.frag:149-188, filter mode 0,
// no clamp) times the pixel BRDF (the commented call at Gather.frag:134).
template <typename T>
__device__ inline V<T> antirad_cone(V<T> pos, V<T> nrm, const Mat<T>& mat, V<T> apos, V<T> aw, V<T> aA,
                                    V<T> cam, T cos_half) {
    const V<T> zero = mk<T>(0, 0, 0);
    const V<T> to_pos = normalize(pos - apos);
    const T dist = length(pos - apos);
    const T c = dot(to_pos, normalize(aw));
    if (c <= T(kConeEps)) return zero;
    const T cp = dot(nrm, -to_pos);
    if (cp <= T(kConeEps)) return zero;
    if (!(c > cos_half)) return zero;  // acos(c) < half-angle
    const T GA = cp / (dist * dist);
    return f_r3(apos, pos, cam, nrm, mat) * (GA * aA);
}

}  // namespace gm
}  // namespace arc
