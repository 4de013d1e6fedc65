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
#include <algorithm>
#include <vector>

#include "arc_internal.hpp"

struct arc_batch {
    arc_ctx* ctx = nullptr;                 // the caller's context: launches run on its stream
    std::vector<arc_ctx*> lanes;            // contexts with streams of their own
    std::vector<hipEvent_t> join;           // one per lane
    hipStream_t origin = nullptr;           // the stream captured from: the caller's may be the null stream
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nviews = 0;
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
        hipEvent_t e = nullptr;
        if (st == ARC_OK) {
            const hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (he != hipSuccess) st = arc::hip_fail(ctx, he, "hipEventCreate", __FILE__, __LINE__);
            else b->join.push_back(e);
        }
    }
    if (st == ARC_OK) {
        const hipError_t he = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
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
    // 1. every view once, eagerly: buffers reach their final sizes.  The
    // lanes first wait for the caller's stream, which produced the inputs.
    ARC_HIP(ctx, hipEventRecord(b->fork, ctx->stream));
    for (arc_ctx* c : b->lanes) ARC_HIP(ctx, hipStreamWaitEvent(c->stream, b->fork, 0));
    for (int i = 0; i < nviews; ++i) {
        const arc_batch_view& v = views[i];
        ARC_TRY(arc_gather_run(v.gather, &v.avpls, v.camPos, v.sceneExtent, &v.params, ARC_TRUST_INPUT, nullptr));
    }
    for (arc_ctx* c : b->lanes) ARC_HIP(ctx, hipStreamSynchronize(c->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // 2. capture on the batch's own stream (the legacy null stream cannot be
    // captured): fork the lanes from it, run the views on their lanes, join
    // back.  The launch orders the whole graph on the caller's stream.
    ARC_HIP(ctx, hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
    int st = ARC_OK;
    hipError_t e = hipEventRecord(b->fork, b->origin);
    for (size_t l = 0; e == hipSuccess && l < b->lanes.size(); ++l) e = hipStreamWaitEvent(b->lanes[l]->stream, b->fork, 0);
    if (e != hipSuccess) st = arc::hip_fail(ctx, e, "batch fork", __FILE__, __LINE__);
    for (int i = 0; st == ARC_OK && i < nviews; ++i) {
        const arc_batch_view& v = views[i];
        st = arc_gather_run(v.gather, &v.avpls, v.camPos, v.sceneExtent, &v.params, ARC_TRUST_INPUT, nullptr);
    }
    for (size_t l = 0; st == ARC_OK && l < b->lanes.size(); ++l) {
        e = hipEventRecord(b->join[l], b->lanes[l]->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[l], 0);
        if (e != hipSuccess) st = arc::hip_fail(ctx, e, "batch join", __FILE__, __LINE__);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->origin, &g);
    if (st == ARC_OK && ee != hipSuccess) st = arc::hip_fail(ctx, ee, "hipStreamEndCapture", __FILE__, __LINE__);
    if (st != ARC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return arc::hip_fail(ctx, e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    b->nviews = nviews;
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
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (arc_ctx* c : b->lanes) arc_ctx_destroy(c);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

}  // extern "C"
