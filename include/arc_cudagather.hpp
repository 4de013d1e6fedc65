/*
 * arc_cudagather.hpp — header-only C++ facade over arc.h with the class and
 * method names of the reference's CUDA path, for host code that called
 * CudaGather (Framework/cudaGather.h:25-70), Bvh / AvplBvh / SimpleBvh
 * (Framework/bvh.h:53-161) and AvplsGpu (Framework/data_gpu.h:14-40).
 *
 * Differences, all forced by the boundary being a C ABI on a GL-free host:
 *  - OpenGL texture names become device buffers (float4 per pixel, row-major);
 *  - CConfigManager becomes arc_gather_params (same field names);
 *  - errors throw arc::Error instead of being ignored;
 *  - the AvplBvh random numbers are an explicit input or a seed instead of
 *    the global std::mt19937 (bvh.cu:27-28).
 * Link with libarc.so.  Only arc.h types appear; no HIP header is needed.
 */
#ifndef ARC_CUDAGATHER_HPP_
#define ARC_CUDAGATHER_HPP_

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "arc.h"

namespace arc {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, arc_ctx* ctx = nullptr) {
    if (status >= 0) return;
    std::string msg = arc_status_string(status);
    if (ctx) {
        const char* e = arc_ctx_last_error(ctx);
        if (e && *e) msg += std::string(": ") + e;
    }
    throw Error(status, msg);
}

/* CudaContext (CudaResources/cudaContext.hpp:15-45): one per host thread. */
class Context {
public:
    explicit Context(int device = 0) { check(arc_ctx_create(device, &ctx_)); }
    ~Context() { arc_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    arc_ctx* get() const { return ctx_; }
    void setStream(void* hipStream) { check(arc_ctx_set_stream(ctx_, hipStream), ctx_); }
    void synchronize() { check(arc_ctx_synchronize(ctx_), ctx_); }

private:
    arc_ctx* ctx_ = nullptr;
};

/* A device allocation owned by the facade (thrust::device_vector's role). */
template <class T>
class DeviceVector {
public:
    DeviceVector() = default;
    DeviceVector(Context& c, size_t n) : ctx_(c.get()), n_(n) {
        if (n) check(arc_device_alloc(ctx_, n * sizeof(T), reinterpret_cast<void**>(&p_)), ctx_);
    }
    DeviceVector(Context& c, const std::vector<T>& host) : DeviceVector(c, host.size()) { upload(host); }
    ~DeviceVector() { release(); }
    DeviceVector(DeviceVector&& o) noexcept { swap(o); }
    DeviceVector& operator=(DeviceVector&& o) noexcept { release(); swap(o); return *this; }
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;

    T* data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const std::vector<T>& host) {
        if (host.size() != n_) throw Error(ARC_E_INVALID, "DeviceVector::upload: size mismatch");
        if (n_) check(arc_copy_to_device(ctx_, p_, host.data(), n_ * sizeof(T)), ctx_);
    }
    std::vector<T> download() const {
        std::vector<T> h(n_);
        if (n_) check(arc_copy_to_host(ctx_, h.data(), p_, n_ * sizeof(T)), ctx_);
        return h;
    }

private:
    void release() {
        if (p_) arc_device_free(ctx_, p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(DeviceVector& o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    arc_ctx* ctx_ = nullptr;
    T* p_ = nullptr;
    size_t n_ = 0;
};

template <class V>
inline arc_float3 to_float3(const V& v) {
    return arc_float3{static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)};
}

/* AvplsGpu (data_gpu.cu:5-52): the AVPL SoA in device memory, packed from any
 * AVPL class with the reference's accessors (Avpl, AVPL.h:34-41):
 * getPosition, getNormal, getIncidentDirection, getIncidentRadiance,
 * getAntiradiance (x/y/z members), getMaterialIndex, getBounce. */
class AvplsGpu {
public:
    template <class AvplT>
    AvplsGpu(Context& c, const std::vector<AvplT>& avpls) : numAvpls(static_cast<int>(avpls.size())) {
        std::vector<arc_float3> pos, nrm, dir, rad, anti;
        std::vector<int> mat, bounce;
        pos.reserve(avpls.size());
        nrm.reserve(avpls.size());
        dir.reserve(avpls.size());
        rad.reserve(avpls.size());
        anti.reserve(avpls.size());
        mat.reserve(avpls.size());
        bounce.reserve(avpls.size());
        for (const AvplT& a : avpls) {
            pos.push_back(to_float3(a.getPosition()));
            nrm.push_back(to_float3(a.getNormal()));
            dir.push_back(to_float3(a.getIncidentDirection()));
            rad.push_back(to_float3(a.getIncidentRadiance()));
            anti.push_back(to_float3(a.getAntiradiance()));
            mat.push_back(a.getMaterialIndex());
            bounce.push_back(a.getBounce());
        }
        fill(c, pos, nrm, dir, rad, anti, mat, bounce);
    }
    AvplsGpu(Context& c, const std::vector<arc_float3>& position, const std::vector<arc_float3>& normal,
             const std::vector<arc_float3>& incDirection, const std::vector<arc_float3>& incRadiance,
             const std::vector<arc_float3>& antiradiance, const std::vector<int>& materialIndex,
             const std::vector<int>& bounce)
        : numAvpls(static_cast<int>(position.size())) {
        fill(c, position, normal, incDirection, incRadiance, antiradiance, materialIndex, bounce);
    }

    /* AvplsGpuParam (data_gpu.h:14-24) with device pointers. */
    arc_avpls param() const {
        return arc_avpls{position.data(), normal.data(), incDirection.data(), incRadiance.data(),
                         antiradiance.data(), materialIndex.data(), bounce.data(), numAvpls};
    }

    int numAvpls;
    DeviceVector<arc_float3> position, normal, incDirection, incRadiance, antiradiance;
    DeviceVector<int> materialIndex, bounce;

private:
    void fill(Context& c, const std::vector<arc_float3>& p, const std::vector<arc_float3>& n,
              const std::vector<arc_float3>& d, const std::vector<arc_float3>& r,
              const std::vector<arc_float3>& a, const std::vector<int>& m, const std::vector<int>& b) {
        const size_t k = p.size();
        if (n.size() != k || d.size() != k || r.size() != k || a.size() != k || m.size() != k || b.size() != k)
            throw Error(ARC_E_INVALID, "AvplsGpu: arrays of different lengths");
        position = DeviceVector<arc_float3>(c, p);
        normal = DeviceVector<arc_float3>(c, n);
        incDirection = DeviceVector<arc_float3>(c, d);
        incRadiance = DeviceVector<arc_float3>(c, r);
        antiradiance = DeviceVector<arc_float3>(c, a);
        materialIndex = DeviceVector<int>(c, m);
        bounce = DeviceVector<int>(c, b);
    }
};

/* Bvh (bvh.h:53-102).  create() runs in the derived constructors, as the
 * reference's derived constructors call it (bvh.cu:773, 876). */
class Bvh {
public:
    virtual ~Bvh() {
        if (owned_) arc_bvh_destroy(bvh_);
    }
    Bvh(const Bvh&) = delete;
    Bvh& operator=(const Bvh&) = delete;

    /* BvhParam (device pointers, bvh.h:72) */
    arc_bvh_param getBvhParam() const {
        arc_bvh_param p{};
        check(arc_bvh_get_param(bvh_, &p), ctx_);
        return p;
    }
    /* BvhData (bvh.h:73) */
    arc_bvh_data getBvhData() const {
        arc_bvh_data d{};
        check(arc_bvh_get_data(bvh_, &d), ctx_);
        return d;
    }
    /* getNode (bvh.cu:660-667) */
    arc_bvh_node getNode(int i) const {
        arc_bvh_node n{};
        check(arc_bvh_get_node(bvh_, i, &n), ctx_);
        return n;
    }
    /* checkTreeIntegrity (bvh.cu:568-639) on the built tree: true when every
     * node is reached once, with two children and matching parents. */
    bool checkTreeIntegrity(int* maxStack = nullptr) const {
        int ms = 0;
        const int st = arc_bvh_check_integrity(bvh_, &ms);
        if (maxStack) *maxStack = ms;
        if (st == ARC_E_STATE) return false;
        check(st, ctx_);
        return true;
    }
    arc_bvh* handle() const { return bvh_; }

protected:
    Bvh(Context& c, bool considerNormals) : ctx_(c.get()) {
        check(arc_bvh_create(ctx_, considerNormals ? 1 : 0, &bvh_), ctx_);
    }
    explicit Bvh(arc_bvh* borrowed, arc_ctx* ctx) : ctx_(ctx), bvh_(borrowed), owned_(false) {}
    arc_ctx* ctx_;
    arc_bvh* bvh_ = nullptr;
    bool owned_ = true;
};

/* SimpleBvh(positions, normals, considerNormals) (bvh.h:150-161) over host
 * vectors. */
class SimpleBvh : public Bvh {
public:
    SimpleBvh(Context& c, const std::vector<arc_float3>& positions, const std::vector<arc_float3>& normals,
              bool considerNormals)
        : Bvh(c, considerNormals) {
        if (!normals.empty() && normals.size() != positions.size())
            throw Error(ARC_E_INVALID, "SimpleBvh: positions and normals differ in length");
        check(arc_bvh_build_points(bvh_, positions.data(), normals.empty() ? nullptr : normals.data(),
                                   static_cast<int>(positions.size()), ARC_INPUT_HOST),
              ctx_);
    }
};

/* AvplBvh(avpls, considerNormals) (bvh.h:129-148) over AVPLs in device
 * memory.  rand: 2n-1 device floats in [0,1) (the reference's
 * randomNumbers), or nullptr for the seeded generator. */
class AvplBvh : public Bvh {
public:
    AvplBvh(Context& c, const AvplsGpu& avpls, bool considerNormals, const float* rand = nullptr,
            uint64_t seed = 5489)
        : Bvh(c, considerNormals) {
        const arc_avpls p = avpls.param();
        check(arc_bvh_build_avpls(bvh_, &p, rand, seed, 0), ctx_);
    }
    /* getAvplBvhNodeDataParam (bvh.h:141), device pointers */
    arc_avpl_node_data getAvplBvhNodeDataParam() const {
        arc_avpl_node_data d{};
        check(arc_bvh_get_node_data(bvh_, &d), ctx_);
        return d;
    }

private:
    friend class CudaGather;
    AvplBvh(arc_bvh* borrowed, arc_ctx* ctx) : Bvh(borrowed, ctx) {}
};

/* CudaGather (cudaGather.h:25-70).  The constructor's GL textures are
 * device buffers: positions and normals (float4, normal.w = material index)
 * and the three float4 outputs (any may be nullptr).  CConfigManager's
 * fields are arc_gather_params. */
class CudaGather {
public:
    CudaGather(Context& c, int width, int height, const float* devPositions4, const float* devNormals4,
               float* devRadianceOutput4, float* devAntiradianceOutput4, float* devResultOutput4,
               const std::vector<arc_material>& materials, const arc_gather_params& conf,
               const float* devAlbedo4 = nullptr)
        : ctx_(c.get()), conf_(conf), width_(width), height_(height) {
        check(arc_gather_create(ctx_, width, height, materials.data(), static_cast<int>(materials.size()), &g_),
              ctx_);
        try {
            check(arc_gather_set_gbuffer(g_, devPositions4, devNormals4, devAlbedo4), ctx_);
            check(arc_gather_set_outputs(g_, devResultOutput4, devRadianceOutput4, devAntiradianceOutput4), ctx_);
        } catch (...) {
            arc_gather_destroy(g_);
            throw;
        }
    }
    ~CudaGather() { arc_gather_destroy(g_); }
    CudaGather(const CudaGather&) = delete;
    CudaGather& operator=(const CudaGather&) = delete;

    /* run(avpls, cameraPosition, sceneProbe, sceneExtent, profile)
     * (cudaGather.h:40-41); the scene probe (a debug pick) is dropped. */
    template <class Vec3>
    arc_gather_stats run(const AvplsGpu& avpls, const Vec3& cameraPosition, float sceneExtent, bool profile) {
        const arc_avpls p = avpls.param();
        arc_gather_stats st{};
        check(arc_gather_run(g_, &p, to_float3(cameraPosition), sceneExtent, &conf_,
                             profile ? ARC_RUN_PROFILE : 0, &st),
              ctx_);
        return st;
    }
    /* Same, packing the host AVPLs first (AvplsGpu per frame, cudaGather.cu:840). */
    template <class AvplT, class Vec3>
    arc_gather_stats run(const std::vector<AvplT>& avpls, const Vec3& cameraPosition, float sceneExtent,
                         bool profile) {
        return runHost(avpls, cameraPosition, sceneExtent, profile);
    }

    arc_gather_params& config() { return conf_; }
    int width() const { return width_; }
    int height() const { return height_; }
    arc_gather* handle() const { return g_; }

    /* The AvplBvh of the last light-cut frame (borrowed view; the reference
     * built it on the stack of run, cudaGather.cu:858). */
    bool hasAvplBvh() const { return arc_gather_get_bvh(g_) != nullptr; }
    AvplBvh getAvplBvh() const {
        arc_bvh* b = arc_gather_get_bvh(g_);
        if (!b) throw Error(ARC_E_STATE, "CudaGather: no light-cut frame has run");
        return AvplBvh(b, ctx_);
    }

private:
    template <class AvplT, class Vec3>
    arc_gather_stats runHost(const std::vector<AvplT>& avpls, const Vec3& cameraPosition, float sceneExtent,
                             bool profile) {
        std::vector<arc_float3> pos, nrm, dir, rad, anti;
        std::vector<int> mat, bounce;
        for (const AvplT& a : avpls) {
            pos.push_back(to_float3(a.getPosition()));
            nrm.push_back(to_float3(a.getNormal()));
            dir.push_back(to_float3(a.getIncidentDirection()));
            rad.push_back(to_float3(a.getIncidentRadiance()));
            anti.push_back(to_float3(a.getAntiradiance()));
            mat.push_back(a.getMaterialIndex());
            bounce.push_back(a.getBounce());
        }
        const arc_avpls p{pos.data(), nrm.data(), dir.data(), rad.data(), anti.data(), mat.data(), bounce.data(),
                          static_cast<int>(avpls.size())};
        arc_gather_stats st{};
        check(arc_gather_run(g_, &p, to_float3(cameraPosition), sceneExtent, &conf_,
                             ARC_INPUT_HOST | (profile ? ARC_RUN_PROFILE : 0), &st),
              ctx_);
        return st;
    }

    arc_ctx* ctx_;
    arc_gather* g_ = nullptr;
    arc_gather_params conf_;
    int width_, height_;
};

/* The reference's defaults (CConfigManager.cpp:14-37). */
inline arc_gather_params defaultConfig() {
    arc_gather_params p;
    arc_gather_params_default(&p);
    return p;
}

}  // namespace arc

#endif  // ARC_CUDAGATHER_HPP_
