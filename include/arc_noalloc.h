/* The allocation-free SimpleBvh build of libarc.
 *
 * arc_bvh_build_points_noalloc has no counterpart in the reference: it is
 * arc_bvh_build_points (Bvh::create over SimpleBvh points, bvh.cu:325-454,
 * bit-exact with it) for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It calls no hipMalloc, hipFree or allocating library routine.
 * The caller passes the inputs (device pointers) and owns the outputs; the
 * build's scratch is a fixed device array of the library, about 18.2 MB per
 * GPU, held by each GPU the library runs on until the process ends.
 *
 * Limits: 2 <= n <= ARC_NOALLOC_MAX_POINTS, larger n returns ARC_E_INVALID.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (two streams, two threads) race and must be serialised by the caller.
 * The work is enqueued on the context's stream; the call does not wait. */
#ifndef ARC_NOALLOC_H_
#define ARC_NOALLOC_H_

#include <stdint.h>

#include "arc.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ARC_NOALLOC_MAX_POINTS (1 << 20)

/* Device buffers the build writes, in the layouts of arc_bvh_data:
 * codes[n], perm[n], nodes[n - 1], parents[2n - 1], ranges[2 (n - 1)],
 * positions[3n] (sorted points) and, when normals are given, normals[3n]. */
typedef struct arc_lbvh_outputs {
    uint32_t* codes;
    int* perm;
    arc_bvh_node* nodes;
    int* parents;
    int* ranges;
    float* positions;
    float* normals;
} arc_lbvh_outputs;

int arc_bvh_build_points_noalloc(arc_ctx* ctx, const arc_float3* positions, const arc_float3* normals, int n,
                                 int consider_normals, const arc_lbvh_outputs* out);

#ifdef __cplusplus
}
#endif

#endif /* ARC_NOALLOC_H_ */
