/*
 * arc.h — C ABI of the MI355X-native AVPL gather and Karras LBVH.
 *
 * Drop-in boundary for the reference's CUDA hot path.  Each declaration names
 * the reference interface it replaces (file:line, flosmn/AntiradianceCuts,
 * Framework/).  Plain pointers and sizes only: device pointers are HIP device
 * addresses on the context's device; host pointers are marked as such.
 *
 * Errors: every entry point returns ARC_OK (0) or a negative ARC_E_* code; the
 * reference ignored errors in release builds (CudaResources/cudaUtil.hpp:32-34)
 * and signalled traversal failures in-band; here both are reported, and the
 * in-band sentinel pixel is kept as an option (arc_gather_params.writeSentinel).
 *
 * Threading: one arc_ctx per host thread.  A context owns a HIP stream (or
 * borrows one, arc_ctx_set_stream); all work of objects made from it is
 * ordered on that stream.  Entry points are asynchronous with respect to the
 * host unless stated otherwise; arc_ctx_synchronize waits.
 */
#ifndef ARC_H_
#define ARC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define ARC_OK 0
#define ARC_E_INVALID (-1)         /* bad argument                                  */
#define ARC_E_TOO_FEW_POINTS (-2)  /* n <= 1: "not enough points" (bvh.cu:327-330)  */
#define ARC_E_HIP (-3)             /* HIP runtime error (see arc_ctx_last_error)    */
#define ARC_E_NOMEM (-4)           /* device allocation failed                      */
#define ARC_E_NO_DEVICE (-5)       /* no HIP device                                 */
#define ARC_E_UNSUPPORTED (-6)     /* combination not implemented                   */
#define ARC_E_STATE (-7)           /* object not ready (e.g. no G-buffer bound)     */

/* ------------------------------------------------------------ POD types */
/* CUDA float3 layout (12 bytes). */
typedef struct arc_float3 { float x, y, z; } arc_float3;

/* MATERIAL (Material.h:6-12), glm::vec3 members: 40 bytes. */
typedef struct arc_material {
    arc_float3 emissive, diffuse, specular;
    float exponent;
} arc_material;

/* BvhNode (bvh.h:21-28): child >= 0 is a leaf (sorted slot), child < 0 is
 * internal node -(child + 1). */
typedef struct arc_bvh_node {
    int left, right;
    arc_float3 bbMin, bbMax;
    int visited;
} arc_bvh_node;

/* AvplsGpuParam (data_gpu.h:14-24): SoA of numAvpls entries. */
typedef struct arc_avpls {
    const arc_float3* position;
    const arc_float3* normal;
    const arc_float3* incDirection;
    const arc_float3* incRadiance;
    const arc_float3* antiradiance;
    const int* materialIndex;
    const int* bounce;
    int numAvpls;
} arc_avpls;

/* BvhParam (bvh.h:44-51): device pointers owned by the arc_bvh. */
typedef struct arc_bvh_param {
    int numLeafs;
    int numNodes;
    arc_bvh_node* nodes;      /* [numNodes]                        */
    arc_float3* positions;    /* [numLeafs] leaf positions, sorted */
    arc_float3* normals;      /* [numLeafs] leaf normals, sorted   */
} arc_bvh_param;

/* BvhData (bvh.h:30-36) plus the outputs the reference keeps transient. */
typedef struct arc_bvh_data {
    uint32_t* morton;         /* [numLeafs] sorted codes (BvhData::morton)            */
    int* parents;             /* [numLeafs + numNodes] leaves then internal nodes     */
    int numLeafs;
    int numNodes;
    int* perm;                /* [numLeafs] input index of each sorted slot           */
    int* ranges;              /* [2 * numNodes] first, last leaf of each node (bvh.cu:222) */
} arc_bvh_data;

/* AvplBvhNodeDataParam (bvh.h:104-113): SoA over 2 * numLeafs - 1 slots,
 * leaves at [0, numLeafs), internal node k at numLeafs + k. */
typedef struct arc_avpl_node_data {
    int* size;
    int* materialIndex;
    float* randomNumbers;
    arc_float3* position;
    arc_float3* normal;
    arc_float3* incRadiance;
    arc_float3* incDirection;
} arc_avpl_node_data;

/* --------------------------------------------------------------- context */
typedef struct arc_ctx arc_ctx;

/* CudaContext (CudaResources/cudaContext.hpp:15-45) picks the device; here
 * the caller names it. */
int arc_ctx_create(int device, arc_ctx** out);
void arc_ctx_destroy(arc_ctx* ctx);
/* Order all work on `hip_stream` (a hipStream_t); NULL is the device's
 * default (null) stream, as used by hosts that never create streams. */
int arc_ctx_set_stream(arc_ctx* ctx, void* hip_stream);
/* Go back to the stream the context created (the state after create). */
int arc_ctx_use_own_stream(arc_ctx* ctx);
void* arc_ctx_get_stream(arc_ctx* ctx);
int arc_ctx_synchronize(arc_ctx* ctx);
const char* arc_ctx_last_error(arc_ctx* ctx);
const char* arc_status_string(int status);
/* Library version string, e.g. "arc 0.1.0 gfx950". */
const char* arc_version(void);

/* Device memory helpers so host code needs no HIP headers (CudaBuffer,
 * CudaResources/cudaBuffer.hpp:11-49).  Copies are synchronous. */
int arc_device_alloc(arc_ctx* ctx, size_t bytes, void** out);
int arc_device_free(arc_ctx* ctx, void* ptr);
int arc_copy_to_device(arc_ctx* ctx, void* dst, const void* host_src, size_t bytes);
int arc_copy_to_host(arc_ctx* ctx, void* host_dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------ LBVH */
typedef struct arc_bvh arc_bvh;

#define ARC_BVH_CONSIDER_NORMALS 1   /* 6D position+normal codes, truncated to 32 bits (bvh.h:17) */
#define ARC_INPUT_HOST 256           /* input pointers are host memory, copied before the call returns */
#define ARC_TRUST_INPUT 512          /* skip the synchronous bounce > 0 check of AVPL inputs      */

/* Bvh(bool considerNormals) (bvh.h:56). */
int arc_bvh_create(arc_ctx* ctx, int consider_normals, arc_bvh** out);
void arc_bvh_destroy(arc_bvh* bvh);

/* SimpleBvh(positions, normals, considerNormals) (bvh.cu:859-877) followed by
 * Bvh::create (bvh.cu:325-454): bounds, normalise, Morton encode, stable key
 * sort, Karras radix tree, bottom-up AABB refit.  `normals` may be NULL when
 * the Bvh does not consider normals.  flags: ARC_INPUT_HOST.
 * Returns ARC_E_TOO_FEW_POINTS when n <= 1. */
int arc_bvh_build_points(arc_bvh* bvh, const arc_float3* positions, const arc_float3* normals,
                         int n, int flags);

/* AvplBvh(avpls, considerNormals) (bvh.cu:746-774): leaves are the AVPLs,
 * node data (AvplBvhNodeData, bvh.cu:673-724) is permuted with the sort and
 * internal nodes get cluster representatives (bvh.cu:726-744) drawn with
 * rand[slot] (2n-1 floats in [0,1), device or host per flags).  When rand is
 * NULL the numbers come from a counter-based generator keyed by `seed`
 * (arc_bvh_random_numbers states it).  Every AVPL must have bounce > 0:
 * the reference builds leaves from bounce > 0 but node data from all AVPLs
 * (bvh.cu:753-759 vs 771); a bounce-0 AVPL returns ARC_E_INVALID. */
int arc_bvh_build_avpls(arc_bvh* bvh, const arc_avpls* avpls, const float* rand, uint64_t seed,
                        int flags);

/* The generator used when rand is NULL: writes 2n-1 floats in [0,1) for
 * `seed` into `out` (device memory). */
int arc_bvh_random_numbers(arc_ctx* ctx, uint64_t seed, int count, float* out);

/* Bvh::getBvhParam (bvh.h:72), getBvhData (bvh.h:73),
 * AvplBvh::getAvplBvhNodeDataParam (bvh.h:141).  Pointers are borrowed and
 * stay valid until the next build or destroy.  Node data is all-NULL for a
 * points-only build. */
int arc_bvh_get_param(arc_bvh* bvh, arc_bvh_param* out);
int arc_bvh_get_data(arc_bvh* bvh, arc_bvh_data* out);
int arc_bvh_get_node_data(arc_bvh* bvh, arc_avpl_node_data* out);

/* Bvh::getNode (bvh.cu:660-667): copies one node to the host (synchronous). */
int arc_bvh_get_node(arc_bvh* bvh, int i, arc_bvh_node* out);

/* Bvh::checkTreeIntegrity (bvh.cu:568-639) on the device arrays (copied to
 * the host, synchronous).  Returns ARC_OK when every node is reached once,
 * every internal node has two children and parents agree; ARC_E_STATE
 * otherwise.  *max_stack receives the largest DFS stack depth. */
int arc_bvh_check_integrity(arc_bvh* bvh, int* max_stack);

/* ---------------------------------------------------------------- gather */
typedef struct arc_gather arc_gather;

#define ARC_CUT_SOLID_ANGLE 0   /* kernel_bvh: Omega < clusterRefinementThreshold (cudaGather.cu:410-414) */
#define ARC_CUT_ERROR_BOUND 1   /* Lightcuts-style: refine while bound > errorBound * estimate           */
#define ARC_CUT_LEARNED 2       /* per 8x8 tile, refine / keep predicted by an arc_cutnet; the error-bound
                                   criterion decides where the prediction is not confident              */

/* What the error-bound cut compares with errorBound * estimate. */
#define ARC_CRITERION_BOUND 0     /* upper bound of the node's contribution (Lightcuts)       */
#define ARC_CRITERION_ESTIMATE 1  /* contribution bound x expected relative error of the node */
#define ARC_CRITERION_EXPANSION 2 /* contribution bound x (s / d)^2, and s <= cutOpening x d, with s the
                                     node's half-diagonal and d the distance to its centroid          */

/* How the error-bound cut sums a kept node. */
#define ARC_ESTIMATOR_CENTROID 0  /* I x G at the node's intensity-weighted centroid, with the mean
                                     normal's clamped-cosine heuristic (and cutEvalDepth sums)       */
#define ARC_ESTIMATOR_MOMENTS 1   /* per RGB channel, the clamped cosine of the node's AVPL normals
                                     to second order in spherical harmonics, times the pixel-side
                                     G averaged over four points that match the node's position
                                     mean and covariance; nodes of at most cutExactCount AVPLs are
                                     summed exactly over their leaves (per 8x8 tile only)          */

#define ARC_ANTIRAD_DISK 0      /* getAntiradiance, cudaGather.cu:109-144                 */
#define ARC_ANTIRAD_CONE 1      /* GetAntiradiance, Shaders/Gather.frag:149-188 x BRDF    */

/* How antiradiance is summed; both give the exact sum over all AVPLs. */
#define ARC_ANTI_EXACT 0        /* every AVPL per pixel (kernel_antiradiance, cudaGather.cu:446-557)     */
#define ARC_ANTI_CLUSTERED 1    /* per screen-cluster AVPL lists (useClusteredDeferred, :605-748, 950-1011) */

/* The CONF_VARS fields the hot path reads (CConfigManager.h:6-52), plus the
 * extension switches.  arc_gather_params_default fills the reference's
 * defaults (CConfigManager.cpp:14-37). */
typedef struct arc_gather_params {
    int useLightCuts;                 /* 1: cut radiance (kernel_bvh), 0: exact (kernel_radiance_simple) */
    int cutMode;                      /* ARC_CUT_*                                   */
    float clusterRefinementThreshold; /* solid-angle threshold, default pi/10        */
    float errorBound;                 /* ARC_CUT_ERROR_BOUND relative bound, 0.01    */
    int antiradianceModel;            /* ARC_ANTIRAD_*                               */
    int antiradianceMode;             /* ARC_ANTI_*                                  */
    float photonRadiusScale;          /* disk radius = scale * sceneExtent / 100     */
    float coneHalfAngle;              /* radians, CONE model                         */
    int writeSentinel;                /* 1: (1e5, 0, 1e5, 1) on traversal error      */
    int stackSize;                    /* traversal stack entries, 32 in kernel_bvh   */
    uint64_t seed;                    /* cluster-representative generator seed       */
    int cutEstimateDepth;             /* ARC_CUT_ERROR_BOUND: depth of the estimate cut */
    int cutBudget;                    /* ARC_CUT_ERROR_BOUND: most nodes one cut evaluates */
    int cutPerTile;                   /* ARC_CUT_ERROR_BOUND: 1 = one cut per 8x8 tile, 0 = per pixel */
    int cutCriterion;                 /* ARC_CUT_ERROR_BOUND: ARC_CRITERION_*                         */
    float cutImageFloor;              /* ARC_CUT_ERROR_BOUND: the threshold is errorBound * max(pixel
                                         estimate, cutImageFloor * image-mean estimate); 0 = per pixel */
    float cutConfidence;              /* ARC_CUT_LEARNED: a prediction p is taken when max(p, 1 - p) >=
                                         cutConfidence (default 0.9), else the error-bound criterion */
    int cutEvalDepth;                 /* ARC_CUT_ERROR_BOUND: a kept node is summed as its descendants
                                         this many levels down (0..4; leaves where the subtree ends) */
    int cutEstimator;                 /* ARC_CUT_ERROR_BOUND: ARC_ESTIMATOR_* (default CENTROID)   */
    int cutExactCount;                /* ARC_ESTIMATOR_MOMENTS: a kept node of at most this many AVPLs
                                         (0..128) is summed exactly over its leaves (default 16)   */
    float cutOpening;                 /* ARC_CRITERION_EXPANSION: largest s / d of a kept node (0.3) */
} arc_gather_params;

void arc_gather_params_default(arc_gather_params* p);

typedef struct arc_gather_stats {
    /* phase times in ms (ARC_RUN_PROFILE), as the CudaTimer prints of run() */
    float msUpload, msBvh, msRadiance, msAntiradiance, msCombine, msTotal;
    /* work counters (ARC_RUN_COUNTERS) */
    long long visitedNodes;           /* inner nodes tested                        */
    long long evaluatedNodes;         /* (pixel, node or leaf) radiance evaluations */
    long long antiVisitedNodes;
    long long antiEvaluatedNodes;
    long long errorPixels;            /* traversal stack overflows (node kept unrefined)   */
    int numLeafs;                     /* AVPLs in the LBVH of this frame            */
    long long budgetHits;             /* ARC_CUT_ERROR_BOUND: nodes kept because the
                                         pixel's cutBudget was spent               */
    long long fallbackNodes;          /* ARC_CUT_LEARNED: (tile, node) decisions left to the
                                         error-bound criterion                      */
    long long testedNodes;            /* error-bound and learned cuts: (pixel, inner node)
                                         evaluations of the error-bound criterion   */
    long long exactLeaves;            /* ARC_ESTIMATOR_MOMENTS: (pixel, AVPL) pairs of kept nodes
                                         summed exactly (also counted in evaluatedNodes)           */
} arc_gather_stats;

#define ARC_RUN_PROFILE 1      /* time each phase with HIP events (synchronises)  */
#define ARC_RUN_COUNTERS 2     /* count visited / evaluated nodes; the error-bound
                                  cut then writes each pixel's evaluated-node count
                                  into the radiance output's w (debug)             */
#define ARC_RUN_TIMELINE 4     /* record phase events without synchronising       */
/* ARC_INPUT_HOST (256): the arc_avpls pointers are host memory, uploaded as
 * AvplsGpu does (data_gpu.cu:5-52). */

/* CudaGather(camera, pos, normal, out x3, materials, ...) (cudaGather.h:28-35,
 * cudaGather.cu:805-829).  Materials are host memory, at most 256 entries. */
int arc_gather_create(arc_ctx* ctx, int width, int height, const arc_material* materials,
                      int num_materials, arc_gather** out);
void arc_gather_destroy(arc_gather* g);

/* The G-buffer textures of the constructor: float4 (x, y, z, 1) positions and
 * float4 (nx, ny, nz, materialIndex) normals, width * height each, row-major;
 * albedo (float4, may be NULL) overrides the pixel material's diffuse colour. */
int arc_gather_set_gbuffer(arc_gather* g, const float* pos4, const float* nrm4, const float* albedo4);

/* The output textures of the constructor, float4 each, any may be NULL:
 * result = radiance - antiradiance (kernel_combine, cudaGather.cu:782-803). */
int arc_gather_set_outputs(arc_gather* g, float* result4, float* radiance4, float* antiradiance4);

/* CudaGather::run (cudaGather.cu:836-1051): radiance gather (exact, or cut
 * over an AvplBvh rebuilt from this frame's AVPLs), antiradiance gather with
 * photonRadius = photonRadiusScale * sceneExtent / 100, combine.
 * flags: ARC_RUN_PROFILE | ARC_RUN_COUNTERS | ARC_INPUT_HOST.  stats may be
 * NULL. */
int arc_gather_run(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                   const arc_gather_params* params, int flags, arc_gather_stats* stats);

/* Exact gather in FP64 (the reference image for error metrics): writes
 * radiance and antiradiance as 3 doubles per pixel (device memory).
 * Not in the reference; the reference's exact kernels are FP32. */
int arc_gather_exact_f64(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                         const arc_gather_params* params, int flags, double* radiance3, double* antiradiance3);

/* Phase times (ms) of the last `frames` runs recorded with ARC_RUN_PROFILE or
 * ARC_RUN_TIMELINE, oldest first, 6 floats per frame: upload, bvh,
 * radiance, antiradiance, combine, total.  At most 64 frames are kept; waits
 * for the newest.  Returns the number of frames written or a negative status.
 * Replaces the CudaTimer prints of run() (cudaGather.cu:845-1049). */
int arc_gather_profile_history(arc_gather* g, int frames, float* out);

/* Durations (ms) of each recorded frame's main kernels, oldest first, 2
 * floats per frame: the radiance traversal / exact kernel and the
 * antiradiance per-pixel kernel (0 when the exact pass summed both), each
 * timed with HIP events on the context's stream right around its launch.
 * Same frames and return value as arc_gather_profile_history. */
int arc_gather_kernel_history(arc_gather* g, int frames, float* out);

/* Where those kernels ran within each frame, 4 floats per frame (ms from the
 * frame's start): radiance kernel start and end, antiradiance kernel start
 * and end.  In a cut frame with clustered antiradiance the two run
 * concurrently on two streams, so their spans overlap.  Same frames and
 * return value as arc_gather_profile_history. */
int arc_gather_kernel_spans(arc_gather* g, int frames, float* out);

/* The AvplBvh of the last cut frame (borrowed; NULL before the first).  After
 * a solid-angle cut frame it carries the reference's node data; after an
 * error-bound or learned cut frame, whose tree sums only what those cuts
 * read, it carries the topology and boxes, and arc_bvh_node_data returns no
 * node data. */
arc_bvh* arc_gather_get_bvh(arc_gather* g);

/* Test accessor, no counterpart in the reference (the reference's node data
 * is read through AvplBvh::getAvplBvhNodeDataParam, bvh.h:140): copies the
 * error-bound cut's node records of the last cut frame to host memory.
 * which = 0: the centroid records, 4 float4 per slot over 2n - 1 slots
 * (leaves, then internal nodes); which = 1: the moment estimator's records,
 * 13 float4 per internal node (ARC_ESTIMATOR_MOMENTS frames only).  *count
 * receives the float4 available; at most max_float4 are copied.  Synchronous. */
#define ARC_RECORDS_CENTROID 0
#define ARC_RECORDS_MOMENTS 1
int arc_gather_copy_cut_records(arc_gather* g, int which, float* host, long long max_float4, long long* count);

/* The context a gather was created on. */
arc_ctx* arc_gather_ctx(arc_gather* g);

/* ------------------------------------------------------------ learned cut */
/* Not in the reference: a refine / keep predictor for the light cut.  For each
 * (8x8 pixel tile, BVH node) the tile traversal computes ARC_CUTNET_FEATURES
 * features (node box extent and distance, normal cone, radiance, the
 * error-bound criterion's terms, the tile's normal, extent and threshold
 * statistics) and a 24-64-64-1 MLP (ReLU, sigmoid) with BF16 weights and FP32
 * accumulation on the matrix cores predicts the probability that the node
 * must be refined.  It is trained on labels made by the exact gather: a node
 * is labelled "refine" when some pixel of the tile sees it with an error,
 * against the node's exact sum, above that pixel's threshold.
 *
 * The weights are a host array of ARC_CUTNET_PARAMS floats: feature mean[24],
 * 1/std[24], W1[32][64] (rows 24..31 zero), b1[64], W2[64][64] (W2[in][out]),
 * b2[64], w3[64], b3, then zero padding. */
typedef struct arc_cutnet arc_cutnet;
#define ARC_CUTNET_FEATURES 24
#define ARC_CUTNET_PARAMS 6400

int arc_cutnet_create(arc_ctx* ctx, arc_cutnet** out);
void arc_cutnet_destroy(arc_cutnet* net);
/* Random initial weights (He-normal, from seed) with the given feature
 * normalisation (host arrays of 24; std entries <= 0 count as 1). */
int arc_cutnet_init(arc_cutnet* net, uint64_t seed, const float* feature_mean, const float* feature_std);
int arc_cutnet_set_weights(arc_cutnet* net, const float* weights, int count);   /* host, count = PARAMS */
int arc_cutnet_get_weights(arc_cutnet* net, float* weights, int count);         /* host, synchronous    */
/* Probabilities of `n` feature rows (device arrays, X row-major n x 24). */
int arc_cutnet_forward(arc_cutnet* net, const float* X, int n, float* prob);
/* Minibatch Adam on the weighted binary cross-entropy (device arrays; w may
 * be NULL for unit weights), `epochs` passes over a seeded permutation of the
 * rows.  loss (host, may be NULL) receives each epoch's mean loss.
 * Synchronous. */
int arc_cutnet_train(arc_cutnet* net, const float* X, const float* y, const float* w, int n, int epochs,
                     int batch, float learning_rate, uint64_t seed, float* loss);
/* The network the gather's ARC_CUT_LEARNED frames use (borrowed; NULL detaches). */
int arc_gather_set_cutnet(arc_gather* g, arc_cutnet* net);
/* No counterpart in the reference (a schedule setting of this library's
 * moment cut).  With chunks_per_tile > 0 the moment cut's traversal hands its
 * kept nodes to a second pass that estimates them on the matrix cores,
 * through a pool of chunks_per_tile chunks of 128 nodes per 8x8 tile (device
 * memory 512 bytes per chunk).  A tile that finds the pool full estimates its
 * nodes in the traversal, so the setting changes the speed and the memory,
 * and the image only by rounding.  0 (the default) estimates every kept node
 * in the traversal, where the estimates fill the issue slots that the
 * traversal's dependent loads leave idle.  Read by the next frame.
 * ARC_E_INVALID for a negative count or more than 256 (4.2 GB at 1080p). */
int arc_gather_set_estimate_pool(arc_gather* g, int chunks_per_tile);
/* No counterpart in the reference (launch shapes of this library's kernels,
 * searched by its autotuner).  cut_stack: the moment cut's traversal stack,
 * in nodes, per 8x8 tile (64, 96 or 128; default 96): the deepest cut a tile
 * can refine to; a tile that needs more keeps the node it cannot refine and
 * is counted in errorPixels.  anti_chunk: the clustered antiradiance's list
 * entries filed into the tiles' LDS sub-lists per pass (768, 1536 or 3072;
 * default 1536): more per pass is fewer passes and more LDS per workgroup.
 * Neither changes the image beyond rounding while errorPixels stays 0.  Read
 * by the next frame.  ARC_E_INVALID for other values. */
int arc_gather_set_launch_shape(arc_gather* g, int cut_stack, int anti_chunk);
/* Training rows of one frame: builds the frame's AvplBvh and thresholds as
 * arc_gather_run does with params (cutMode is ignored), then walks each tile's
 * cut with exact labels (a node is refined exactly when its label is
 * "refine") and writes the features and label of each internal node it
 * meets, kept with probability `keep`, into X (capacity x 24) and y
 * (capacity), device memory.  *count receives the number of rows met, which
 * may exceed capacity (the rest are dropped).  Synchronous. */
int arc_gather_cut_samples(arc_gather* g, const arc_avpls* avpls, arc_float3 camPos, float sceneExtent,
                           const arc_gather_params* params, int flags, float* X, float* y, int capacity,
                           uint64_t seed, float keep, int* count);

/* -------------------------------------------------------- batched views */
/* Not in the reference (one view per CudaGather::run on the default stream):
 * many views recorded once into one HIP graph, replayed with one launch.
 * Views are dealt over `lanes` streams of the batch's own, forked from and
 * joined back to ctx's stream, so views on different lanes run together. */
typedef struct arc_batch arc_batch;

typedef struct arc_batch_view {
    arc_gather* gather;          /* created on a lane context (arc_batch_lane), G-buffer and
                                    outputs bound                                              */
    arc_avpls avpls;             /* device AVPLs of the view                                  */
    arc_float3 camPos;
    float sceneExtent;
    arc_gather_params params;
} arc_batch_view;

int arc_batch_create(arc_ctx* ctx, int lanes, arc_batch** out);   /* lanes: 1..8 */
/* The context of lane `lane`, to create that lane's gathers on. */
arc_ctx* arc_batch_lane(arc_batch* b, int lane);
/* Runs every view once (its buffers reach their sizes), then records all of
 * them into one graph.  Synchronous.  The views' inputs and outputs must stay
 * where they are while the graph is in use; their contents may change.  Each
 * view needs a gather of its own (ARC_E_INVALID when two views share one). */
int arc_batch_capture(arc_batch* b, const arc_batch_view* views, int nviews);
/* When every view's radiance is the moment cut (ARC_ESTIMATOR_MOMENTS), the
 * captured graph runs the views' cuts as one launch over all views' tiles,
 * heaviest first (on = 1), or one launch per view (on = 0, the default).  No
 * counterpart in the reference, which renders one view per run; read at the
 * next arc_batch_capture. */
int arc_batch_set_one_cut_launch(arc_batch* b, int on);
/* One graph launch on ctx's stream: a frame of every view.  Asynchronous. */
int arc_batch_launch(arc_batch* b);
void arc_batch_destroy(arc_batch* b);

/* ------------------------------------------------------------ ray tracing */
/* The triangle scene the AVPLs are shot into: Scene with its kd-tree
 * (Scene.cpp:175-255, KdTreeAccelerator.h:17-59).  Triangles follow Triangle
 * (Triangle.h): vertices p0, p1, p2, normal normalize(cross(p2 - p0, p1 - p0)),
 * a front face seen against the normal; a two-sided surface is two triangles.
 * Hits on triangles whose material is emissive count as misses
 * (Scene.cpp:247-252). */
typedef struct arc_trace_scene arc_trace_scene;

/* AreaLight (AreaLight.h): the emitting rectangle is centred at `center`,
 * faces `front`, has its height along `up` projected into its plane.  (The
 * reference places the quad at centerPosition + normalize(frontDirection);
 * pass that point as center.) */
typedef struct arc_area_light {
    arc_float3 center, front, up;
    float width, height;
    arc_float3 radiance;
    int materialIndex;
} arc_area_light;

#define ARC_SHOOT_NO_EXPLICIT_DIRECT 1 /* keep the antiradiance of bounce-1 AVPLs
                                           (explicitDirectIllum off, AvplShooter.h:115-117) */

/* Builds the scene's triangle BVH (flags: ARC_INPUT_HOST).  Materials are
 * host memory, at most 256 entries. */
int arc_trace_scene_create(arc_ctx* ctx, const arc_float3* p0, const arc_float3* p1, const arc_float3* p2,
                           const int* material, int ntri, const arc_material* materials, int num_materials,
                           int flags, arc_trace_scene** out);
void arc_trace_scene_destroy(arc_trace_scene* scene);
/* Triangle count and depth of the scene's BVH. */
int arc_trace_scene_info(arc_trace_scene* scene, int* num_triangles, int* tree_depth);

/* Scene::IntersectRayScene(ray, &t, &isect, FRONT_FACE) for n rays (device
 * arrays): the closest front-face hit with t > 0 (t = -1 and triangle -1 when
 * none or emissive).  Triangle indices are input indices. */
int arc_trace_intersect(arc_trace_scene* scene, const arc_float3* org, const arc_float3* dir, int n, float* t,
                        int* tri_index);

/* AvplShooter::shoot(direct, indirect, numPaths) (AvplShooter.h:36-142): one
 * path per index from the area light, bounces with Russian roulette 0.8.
 * Writes the indirect AVPLs (bounce >= 1, what the gather takes,
 * Render.cpp:297) into device buffers owned by the scene, path by path in
 * index order, and points *out at them (valid until the next shoot or
 * destroy).  Random numbers: a counter-based hash of (seed, path, draw).
 * Synchronous (the AVPL count is read back). */
int arc_trace_shoot(arc_trace_scene* scene, const arc_area_light* light, int num_paths, uint64_t seed, int flags,
                    arc_avpls* out, int* num_indirect);

/* The G-buffer (Shaders/CreateGBuffer.frag) by ray casting: pixel (px, py),
 * row 0 at the top, looks along normalize(F + sx * t * aspect * R + sy * t * U)
 * with F = normalize(forward), R = normalize(cross(up, F)), U = cross(F, R),
 * t = tan(fovy / 2), sx = 2 (px + 0.5) / width - 1, sy = 1 - 2 (py + 0.5) /
 * height.  Writes float4 (x, y, z, 1) positions and (n, material) normals of
 * the closest front face, (0,0,0,0) / (0,0,-1,1) where nothing is hit. */
int arc_trace_gbuffer(arc_trace_scene* scene, arc_float3 camPos, arc_float3 forward, arc_float3 up, float fovy,
                      int width, int height, float* pos4, float* nrm4);

/* ------------------------------------------------------------ output side */
/* Renderer::Add(target, source) (Render.cpp:788-806, Shaders/Add.frag):
 * accum.rgb += frame.rgb, accum.w += 1 (float4 images of npix pixels, device
 * memory).  The reference adds the result texture of each gathered frame. */
int arc_image_accumulate(arc_ctx* ctx, float* accum4, const float* frame4, int npix);

/* Renderer::Normalize (Render.cpp:642-667, Shaders/Normalize.frag):
 * out.rgb = accum.rgb * (1.f / num_paths) (0 when num_paths == 0), out.w = 1. */
int arc_image_normalize(arc_ctx* ctx, float* out4, const float* accum4, int npix, int num_paths);

typedef struct arc_image_error_t {
    float referenceError;  /* CReferenceImage::GetError (CReferenceImage.cpp:28-51):
                              (e.x * e.y + e.z) / 3 with e the per-channel root mean of
                              min(100, d^2), the reference's formula as written        */
    float rmse[3];         /* that e per channel                                        */
    double relL2;          /* |img - ref| / |ref| over RGB                               */
    double maxAbs;         /* max |img - ref| over RGB                                   */
} arc_image_error_t;

/* Error of img4 against ref4 (device float4 images).  Synchronous. */
int arc_image_error(arc_ctx* ctx, const float* ref4, const float* img4, int npix, arc_image_error_t* out);

/* CExport::ExportPFM (Utils/CExport.cpp:22-65): "PF\n<w> <h>\n-1.0\n" and the RGB
 * floats of a host float4 image in memory row order.  Host only. */
int arc_write_pfm(const char* path, const float* rgba_host, int width, int height);
/* Reads such a file into a host float4 image (w = 1).  With rgba_host NULL
 * only the size is returned. */
int arc_read_pfm(const char* path, float* rgba_host, int max_pixels, int* width, int* height);

#ifdef __cplusplus
}
#endif

#endif /* ARC_H_ */
