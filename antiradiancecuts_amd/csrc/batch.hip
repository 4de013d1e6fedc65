// Batched frames on gfx950: many views (each its own G-buffer, AVPL set and
// parameters) recorded once into one HIP graph and replayed with a single
// launch per batch.
//
// The reference renders one view per CudaGather::run call from one host
// thread on the default stream (cudaGather.cu:836-1051).  Small views leave
// most of the 256 CUs idle and every run pays its launch latencies, so the
// batch deals its views round robin over a few lanes (streams of their own,
// forked from the caller's stream and joined back), captures every view's
// complete frame -- LBVH rebuild, cut, antiradiance, combine -- into one
// graph, and replays it with hipGraphLaunch.  Views on different lanes run
// concurrently; the graph removes the per-kernel launch overhead.
//
// Capture needs frames that allocate nothing and never wait on the host: the
// capture call first runs every view once eagerly (buffers reach their final
// sizes), and the captured runs use ARC_TRUST_INPUT with no profiling or
// counters (no host read-backs, no memset nodes).
//
// When every view's radiance is the moment cut, the views' cuts -- most of
// each frame's time -- run as one launch over all views' tiles: each view's
// frame runs up to its cut on its lane, the lanes join, one kernel takes
// the tiles of all views heaviest first (by their costs in the last
// launch), and the lanes fork again for each view's combine.  A small view
// then adds its tiles to the launch instead of leaving the GPU to a short
// kernel of its own, and no view's heaviest tiles run alone at the end.
#include <algorithm>
#include <vector>

#include "gather_common.hpp"

struct arc_batch {
    arc_ctx* ctx = nullptr;                 // the caller's context: launches run on its stream
    std::vector<arc_ctx*> lanes;            // contexts with streams of their own
    std::vector<hipEvent_t> join, join2;    // one per lane each
    hipStream_t origin = nullptr;           // the stream captured from: the caller's may be the null stream
    hipEvent_t fork = nullptr, fork2 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nviews = 0;
    // one launch of the views' moment cuts when every view has one (off by
    // default: on the 64-view workload a cut per view measured faster, as
    // each view's cut then overlaps other views' rebuilds and lists)
    int one_cut = 0;
    arc::DevBuf views_dev, costs, order;   // the views' descriptors, tile costs and launch order
    int nitems = 0;                        // tiles of all views
};

namespace {

constexpr int kMaxLanes = 8;

void drop_graph(arc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->nviews = 0;
}

// Every view's radiance is the moment cut (per 8x8 tile) over > 1 AVPLs.
bool all_moment_cuts(const arc_batch_view* views, int nviews) {
    for (int i = 0; i < nviews; ++i) {
        const arc_gather_params& p = views[i].params;
        if (!(p.useLightCuts && p.cutMode == ARC_CUT_ERROR_BOUND && p.cutEstimator == ARC_ESTIMATOR_MOMENTS &&
              views[i].avpls.numAvpls > 1))
            return false;
    }
    return true;
}

// A view's frame up to its cut (the cut joins the batch's launch).
int run_deferred(const arc_batch_view& v) {
    v.gather->defer_cut = true;
    const int st = arc_gather_run(v.gather, &v.avpls, v.camPos, v.sceneExtent, &v.params, ARC_TRUST_INPUT, nullptr);
    v.gather->defer_cut = false;
    return st;
}

// The views' lanes wait for `ev` recorded on `s`.
int fork_lanes(arc_batch* b, hipStream_t s, hipEvent_t ev) {
    hipError_t e = hipEventRecord(ev, s);
    for (size_t l = 0; e == hipSuccess && l < b->lanes.size(); ++l) e = hipStreamWaitEvent(b->lanes[l]->stream, ev, 0);
    return e == hipSuccess ? ARC_OK : arc::hip_fail(b->ctx, e, "batch fork", __FILE__, __LINE__);
}

// `s` waits for every lane (events `evs`, one per lane).
int join_lanes(arc_batch* b, hipStream_t s, std::vector<hipEvent_t>& evs) {
    hipError_t e = hipSuccess;
    for (size_t l = 0; e == hipSuccess && l < b->lanes.size(); ++l) {
        e = hipEventRecord(evs[l], b->lanes[l]->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, evs[l], 0);
    }
    return e == hipSuccess ? ARC_OK : arc::hip_fail(b->ctx, e, "batch join", __FILE__, __LINE__);
}

// Eagerly, before the capture: every view's frame up to its cut, the views'
// descriptors into device memory, one unordered launch of all cuts (it
// writes the tile costs the captured launches order by), the combines.
int prepare_one_cut(arc_batch* b, const arc_batch_view* views, int nviews) {
    arc_ctx* ctx = b->ctx;
    ARC_TRY(fork_lanes(b, ctx->stream, b->fork));
    for (int i = 0; i < nviews; ++i) ARC_TRY(run_deferred(views[i]));
    for (arc_ctx* c : b->lanes) ARC_HIP(ctx, hipStreamSynchronize(c->stream));
    std::vector<arc::CutView> hv(nviews);
    int nitems = 0;
    for (int i = 0; i < nviews; ++i) {
        hv[i] = views[i].gather->def_view;
        hv[i].tile_base = nitems;
        nitems += hv[i].ntiles;
    }
    ARC_TRY(arc::ensure(ctx, b->views_dev, sizeof(arc::CutView) * (size_t)nviews));
    ARC_TRY(arc::ensure(ctx, b->costs, sizeof(float) * (size_t)nitems));
    ARC_TRY(arc::ensure(ctx, b->order, sizeof(int) * (size_t)nitems));
    for (int i = 0; i < nviews; ++i) hv[i].tile_cost = b->costs.as<float>() + hv[i].tile_base;
    ARC_HIP(ctx, hipMemcpy(b->views_dev.p, hv.data(), sizeof(arc::CutView) * (size_t)nviews, hipMemcpyHostToDevice));
    ARC_TRY(arc::cut_moments_views(ctx, b->origin, b->views_dev.as<arc::CutView>(), nviews, nitems,
                                   b->costs.as<float>(), b->order.as<int>(), false));
    ARC_HIP(ctx, hipStreamSynchronize(b->origin));
    for (int i = 0; i < nviews; ++i) ARC_TRY(arc::gather_finish(views[i].gather));
    for (arc_ctx* c : b->lanes) ARC_HIP(ctx, hipStreamSynchronize(c->stream));
    b->nitems = nitems;
    return ARC_OK;
}

}  // namespace

extern "C" {

int arc_batch_create(arc_ctx* ctx, int lanes, arc_batch** out) {
    if (!ctx || !out || lanes < 1 || lanes > kMaxLanes) return ARC_E_INVALID;
    *out = nullptr;
    arc_batch* b = new arc_batch();
    b->ctx = ctx;
    int st = ARC_OK;
    for (int i = 0; st == ARC_OK && i < lanes; ++i) {
        arc_ctx* c = nullptr;
        st = arc_ctx_create(ctx->device, &c);
        if (st == ARC_OK) b->lanes.push_back(c);
        for (int k = 0; k < 2 && st == ARC_OK; ++k) {
            hipEvent_t e = nullptr;
            const hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (he != hipSuccess) st = arc::hip_fail(ctx, he, "hipEventCreate", __FILE__, __LINE__);
            else (k == 0 ? b->join : b->join2).push_back(e);
        }
    }
    if (st == ARC_OK) {
        hipError_t he = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming);
        if (he != hipSuccess) st = arc::hip_fail(ctx, he, "hipEventCreate", __FILE__, __LINE__);
    }
    if (st == ARC_OK) {
        const hipError_t he = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
        if (he != hipSuccess) st = arc::hip_fail(ctx, he, "hipStreamCreate", __FILE__, __LINE__);
    }
    if (st != ARC_OK) {
        arc_batch_destroy(b);
        return st;
    }
    *out = b;
    return ARC_OK;
}

arc_ctx* arc_batch_lane(arc_batch* b, int lane) {
    if (!b || lane < 0 || lane >= (int)b->lanes.size()) return nullptr;
    return b->lanes[lane];
}

int arc_batch_capture(arc_batch* b, const arc_batch_view* views, int nviews) {
    if (!b || nviews < 1 || !views) return ARC_E_INVALID;
    arc_ctx* ctx = b->ctx;
    for (int i = 0; i < nviews; ++i) {
        if (!views[i].gather) return arc::fail(ctx, ARC_E_INVALID, "batch view without a gather");
        const arc_ctx* vc = arc_gather_ctx(views[i].gather);
        if (std::find(b->lanes.begin(), b->lanes.end(), vc) == b->lanes.end())
            return arc::fail(ctx, ARC_E_INVALID, "a batch view's gather must be created on a lane context");
        // a gather writes its own outputs: two views of one gather would
        // overwrite each other's result inside the graph
        for (int j = 0; j < i; ++j)
            if (views[j].gather == views[i].gather)
                return arc::fail(ctx, ARC_E_INVALID, "two batch views share one gather (each view needs its own)");
    }
    drop_graph(b);
    const bool one_cut = b->one_cut && all_moment_cuts(views, nviews);
    // 1. every view once, eagerly: buffers reach their final sizes.  The
    // lanes first wait for the caller's stream, which produced the inputs.
    ARC_TRY(fork_lanes(b, ctx->stream, b->fork));
    for (int i = 0; i < nviews; ++i) {
        const arc_batch_view& v = views[i];
        ARC_TRY(arc_gather_run(v.gather, &v.avpls, v.camPos, v.sceneExtent, &v.params, ARC_TRUST_INPUT, nullptr));
    }
    for (arc_ctx* c : b->lanes) ARC_HIP(ctx, hipStreamSynchronize(c->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (one_cut) ARC_TRY(prepare_one_cut(b, views, nviews));
    // 2. capture on the batch's own stream (the legacy null stream cannot be
    // captured): fork the lanes from it, run the views on their lanes, join
    // back.  The launch orders the whole graph on the caller's stream.
    ARC_HIP(ctx, hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
    int st = fork_lanes(b, b->origin, b->fork);
    if (!one_cut) {
        for (int i = 0; st == ARC_OK && i < nviews; ++i) {
            const arc_batch_view& v = views[i];
            st = arc_gather_run(v.gather, &v.avpls, v.camPos, v.sceneExtent, &v.params, ARC_TRUST_INPUT, nullptr);
        }
        if (st == ARC_OK) st = join_lanes(b, b->origin, b->join);
    } else {
        // each view up to its cut on its lane; all cuts in one launch; the combines
        for (int i = 0; st == ARC_OK && i < nviews; ++i) st = run_deferred(views[i]);
        if (st == ARC_OK) st = join_lanes(b, b->origin, b->join);
        if (st == ARC_OK)
            st = arc::cut_moments_views(ctx, b->origin, b->views_dev.as<arc::CutView>(), nviews, b->nitems,
                                        b->costs.as<float>(), b->order.as<int>(), true);
        if (st == ARC_OK) st = fork_lanes(b, b->origin, b->fork2);
        for (int i = 0; st == ARC_OK && i < nviews; ++i) st = arc::gather_finish(views[i].gather);
        if (st == ARC_OK) st = join_lanes(b, b->origin, b->join2);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->origin, &g);
    if (st == ARC_OK && ee != hipSuccess) st = arc::hip_fail(ctx, ee, "hipStreamEndCapture", __FILE__, __LINE__);
    if (st != ARC_OK) {
        if (g) (void)hipGraphDestroy(g);
        for (int i = 0; i < nviews; ++i) views[i].gather->deferred = false;
        return st;
    }
    b->graph = g;
    const hipError_t e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return arc::hip_fail(ctx, e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    b->nviews = nviews;
    return ARC_OK;
}

int arc_batch_set_one_cut_launch(arc_batch* b, int on) {
    if (!b || (on != 0 && on != 1)) return ARC_E_INVALID;
    b->one_cut = on;
    return ARC_OK;
}

int arc_batch_launch(arc_batch* b) {
    if (!b) return ARC_E_INVALID;
    if (!b->exec) return arc::fail(b->ctx, ARC_E_STATE, "batch not captured");
    ARC_HIP(b->ctx, hipGraphLaunch(b->exec, b->ctx->stream));
    return ARC_OK;
}

void arc_batch_destroy(arc_batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->join2) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->fork2) (void)hipEventDestroy(b->fork2);
    arc::release(b->views_dev);
    arc::release(b->costs);
    arc::release(b->order);
    for (arc_ctx* c : b->lanes) arc_ctx_destroy(c);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

}  // extern "C"
