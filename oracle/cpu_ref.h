/*
 * ORACLE — test infrastructure, not product code.
 *
 * CPU restatement of the reference's wide-BVH traversal hot path, used only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product path (librt.so) never links, loads or calls it.
 *
 * It consumes the reference FastNode layout (see include/rt.h rt_bvh_desc)
 * exactly as FastNodeManager<G,M,WG> stores it, and reproduces:
 *   aabbIntersect{k}   src/ISPC/rayTracerTemplate.ispc:170-234
 *   triIntersect{k}    rayTracerTemplate.ispc:25-108 (gang-G tie rule)
 *   triAnyHit{k}       rayTracerTemplate.ispc:111-168
 *   intersect          src/accelerationStructure/fastNodeManager.h:1328-1418
 *   intersectSaveDistance                   fastNodeManager.h:1232-1326
 *   intersectSecondary                      fastNodeManager.h:1420-1493
 *   intersectWideAlternative                fastNodeManager.h:835-1049
 *   intersectSecondaryWideAlternative       fastNodeManager.h:1051-1230
 *   getNodeLoopCount / getLeafLoopCount / callIspcTemplateNotConst
 *                                           fastNodeManager.h:1728-1763,271-275
 * plus per-ray node/leaf counters (count definitions of
 * compactNode.cpp:651-717), an FP64 re-evaluation of Moeller-Trumbore and an
 * FP32 brute-force closest hit.
 *
 * Parity pinning (what checks each function; see tests/):
 *   aabb_intersect  closed-form slab cases (test_oracle_kernels.py)
 *   tri_intersect   closed-form MT cases, gang-4/8 tie cases
 *   tri_anyhit      closed-form MT cases incl. the NaN-t acceptance
 *   intersect       brute force on small scenes, hand-built 2-level BVHs
 *                   with known visit sequences and counts, N/L invariance
 *   save_distance   same hit as intersect (brute force), fewer/equal visits
 *   secondary       brute-force any-hit, slot-order push on hand-built BVHs
 *   wide / secondary_wide   per-ray == stream (emulation flags off)
 *   visit_sequences length == nodes + leaves of the counts, root first,
 *                   every later id a child of an earlier visited node, the
 *                   hit leaf visited; hand-built BVH with a known sequence
 */
#ifndef BVHRT_ORACLE_CPU_REF_H
#define BVHRT_ORACLE_CPU_REF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ray {
    float tmax, ox, oy, oz, ix, iy, iz, dx, dy, dz; /* FastRay layout */
} orc_ray;

typedef struct orc_bvh {
    uint32_t M, Lm, G, node_count;
    const float* bounds;
    const uint32_t* base;
    const int8_t* order;
    const uint16_t* child_type;
    uint64_t tri_float_count;
    const float* tri;
} orc_bvh;

typedef struct orc_counts {
    uint32_t nodes, leaves, node_slots, leaf_slots;
} orc_counts;

/* diagnostics per ray */
#define ORC_DIAG_STACK48 1u  /* more than 48 stack entries (reference UB)   */
#define ORC_DIAG_BOGUS 2u    /* pushed a node id >= node_count (reference UB) */
#define ORC_DIAG_STACKCAP 4u /* exceeded the oracle's own stack capacity     */

typedef struct orc_hit {
    int32_t slot;  /* triIndex, -1 = miss */
    uint32_t leaf; /* primIdBegin          */
    float t, b1, b2;
} orc_hit;

/* modes */
#define ORC_CLOSEST 0
#define ORC_ANY 1
#define ORC_SAVE_DIST 2

/* emulation flags for the stream form */
#define ORC_WIDE_AXIS_BUG 1u   /* rayMajorAxis[i] instead of [rayId] (:913,939) */
#define ORC_WIDE_SHARED_DIST 2u/* one aabbDistances array for the whole stream  */
#define ORC_WIDE_ROOT_AS_NODE 4u/* stack[0]=0 is always taken as an inner node  */

/* kernels (exported for the known-answer tests) */
int orc_aabb_intersect(const float* bounds, int k, float* dist, const float* ray_info);
int orc_tri_intersect(const float* tri, uint32_t tri_id, int k, float* ray_info, int gang);
int orc_tri_anyhit(const float* tri, uint32_t tri_id, int k, const float* ray_info);
int orc_node_loop_count(const orc_bvh* b, uint32_t node);
int orc_leaf_loop_count(const orc_bvh* b, uint32_t prim_id_begin);
int orc_dispatch_k(int k);

/* per-ray traversals; return hit/occluded; ray->tmax is updated */
int orc_intersect(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index, int32_t* tri_index,
                  orc_counts* counts, uint32_t* diag);
int orc_intersect_save_distance(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index,
                                int32_t* tri_index, orc_counts* counts, uint32_t* diag);
int orc_intersect_secondary(const orc_bvh* b, orc_ray* ray, orc_counts* counts, uint32_t* diag);

/* visit sequences: per ray, the FastNode ids it visits in order under the
 * per-ray traversal of `mode` (ORC_*), leaves with bit 31 set; ids is
 * n x cap, lens[i] the visits of ray i (the first cap are kept) */
void orc_visit_sequences(const orc_bvh* b, const orc_ray* rays, uint32_t n, int mode, uint32_t* ids,
                         uint32_t cap, uint32_t* lens);

/* stream forms over `count` rays (count <= 4096) */
void orc_intersect_wide(const orc_bvh* b, orc_ray* rays, uint32_t count, uint32_t flags,
                        uint32_t* leaf_index, int32_t* tri_index, orc_counts* counts, uint32_t* diag);
void orc_intersect_secondary_wide(const orc_bvh* b, orc_ray* rays, uint32_t count, uint32_t flags,
                                  uint16_t* result, orc_counts* counts, uint32_t* diag);

/* batch driver: OpenMP dynamic(4) over tiles of `tile` rays like
 * CameraFast::renderImage (cameraFast.h:318-320,462-465).  mode ORC_*;
 * stream != 0 uses the stream traversal per tile.  hits (closest/save) or
 * occluded (any) per ray; counts/diag optional.  threads 0 = OpenMP default.
 * Returns the number of threads used. */
int orc_trace(const orc_bvh* b, const orc_ray* rays, uint32_t n, int mode, int stream,
              uint32_t tile, int threads, orc_hit* hits, uint8_t* occluded, orc_counts* counts,
              uint32_t* diag);

/* FP64 Moeller-Trumbore of one ray against one triangle (9 floats).
 * Returns 1 when the double-precision test accepts (t >= 0, no tMax). */
int orc_mt_f64(const orc_ray* ray, const float* tri9, double* t, double* b1, double* b2);

/* FP32 brute-force closest hit over an unstructured triangle list using the
 * triIntersect arithmetic; returns triangle id or -1, writes t. */
int32_t orc_bruteforce_closest(const float* tris, uint32_t n_tris, const orc_ray* ray, float* t);
int orc_bruteforce_any(const float* tris, uint32_t n_tris, const orc_ray* ray);

#ifdef __cplusplus
}
#endif
#endif
