/*
 * rt.h — C ABI of the MI355X wide-BVH ray-query engine (librt.so).
 *
 * The reference (Hengoo/BVHRaytracer) crosses a C ABI once per node visit and
 * once per leaf visit: the ISPC exports declared in the generated header
 * src/ISPC/ISPCBuild/rayTracer_ISPC.h (included at
 * src/accelerationStructure/fastNodeManager.h:256), i.e.
 *   bool    aabbIntersect{4,8,12,16}(const float* bounds, float* dist, float* rayInfo)
 *   int32_t triIntersect{4,8,12,16}(const float* tri, uint32_t triId, float* rayInfo)
 *   bool    triAnyHit{4,8,12,16}(const float* tri, uint32_t triId, float* rayInfo)
 * (src/ISPC/rayTracerTemplate.ispc:26,112,171), driven by the header-only
 * FastNodeManager traversals (fastNodeManager.h:1232-1493, 835-1230).
 * This library moves that seam to one call per ray batch: the caller hands
 * over the flattened FastNode arrays once (rt_upload_bvh) and then whole ray
 * buffers (rt_trace).  All functions return RT_OK (0) or a negative RT_ERR_*
 * code; rt_last_error() returns the message of the calling thread's last
 * failure.  No C++ exception crosses this boundary.
 *
 * Host-side entry points (rt_scene_*, rt_build*, rt_camera_*, rt_spawn_*,
 * rt_sah_factor) restate the reference's host code that produces the path's
 * inputs: the SAH builder, the FastNode flattener, the camera and the
 * sahFactorConfigSse.txt reader.  The renderer driver above them
 * (config.txt, the N/L sweep, CameraFast, the reports) is the C++ layer in
 * rt_renderer.hpp.
 */
#ifndef BVHRT_RT_H
#define BVHRT_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define RT_OK 0
#define RT_ERR_ARG (-1)         /* bad argument / inconsistent sizes          */
#define RT_ERR_HIP (-2)         /* HIP runtime call failed                    */
#define RT_ERR_NOMEM (-3)       /* device or host allocation failed           */
#define RT_ERR_UNSUPPORTED (-4) /* configuration outside the supported set    */
#define RT_ERR_LAYOUT (-5)      /* FastNode arrays violate the layout contract */
#define RT_ERR_BUSY (-6)        /* serving queue full / admission refused     */

/* ---- ray / hit records ---------------------------------------------------- */

/* == FastRay (src/fastRay.h:18-23): tMax, pos, invDirection, direction.
 * The ISPC kernels read exactly this memory as rayInfo[0..9]
 * (rayTracerTemplate.ispc:46-47,190-192).  dir is NOT normalised. */
typedef struct rt_ray {
    float tmax;
    float ox, oy, oz;
    float ix, iy, iz;
    float dx, dy, dz;
} rt_ray;

/* Closest-hit result.  leaf/slot are the reference's hit identity
 * (leafIndex = FastNode::primIdBegin, triIndex = in-leaf index,
 * fastNodeManager.h:1369-1374); tri maps it back to the input triangle. */
typedef struct rt_hit {
    int32_t tri;   /* input triangle id, -1 = miss                          */
    uint32_t leaf; /* primIdBegin of the hit leaf (float offset), 0 on miss  */
    int32_t slot;  /* in-leaf index, -1 = miss                               */
    float t;       /* final ray.tMax (the closest t)                         */
    float b1, b2;  /* barycentrics of the hit triangle, same FP32 arithmetic */
} rt_hit;

/* Per-ray work counters.  nodes = number of aabbIntersect calls (inner node
 * pops), leaves = number of triIntersect / triAnyHit calls (leaf pops), the
 * count definitions of compactNode.cpp:651-717 applied to the FastNode
 * traversal.  node_slots / leaf_slots = sum of the loop counts k passed to
 * those calls (used for the bytes-per-ray roofline). */
typedef struct rt_counts {
    uint32_t nodes, leaves;
    uint32_t node_slots, leaf_slots;
} rt_counts;

/* ---- the flattened BVH (reference FastNode layout) ------------------------ */

/* Borrowed views of FastNodeManager's arrays (fastNodeManager.h:29-90,160-164,
 * built at :1495-1673): per node i
 *   bounds[i*6*M .. +6M)      SoA child boxes, plane stride nodeRealSize,
 *                             NaN marker at [6M-1] / count at [6M-2];
 *                             bounds[i*6*M] is NaN for a leaf
 *   base[i]                   childIdBegin (inner) / primIdBegin (leaf)
 *   order[i*3*M + axis*M + p] traverseOrderEachAxis, zero padded to M
 *   child_type[i]             childType bitset (bit j = child j is inner)
 * and tri_points = trianglePoints (blocks of 9*leaf_memory floats per leaf,
 * plane stride leafRealSize, NaN marker at [9Lm-1]).  tri_ids maps
 * (leaf ordinal * leaf_memory + slot) to the input triangle id, -1 = pad. */
typedef struct rt_bvh_desc {
    uint32_t branching;   /* N  */
    uint32_t leaf_size;   /* L  */
    uint32_t gang;        /* G  (4 = sse4-i32x4, 8 = avx2-i32x8, 16)          */
    uint32_t node_memory; /* M  = roundup(N, G) */
    uint32_t leaf_memory; /* Lm = roundup(L, G) */
    uint32_t node_count;
    const float* bounds;
    const uint32_t* base;
    const int8_t* order;
    const uint16_t* child_type;
    uint64_t tri_float_count;
    const float* tri_points;
    const int32_t* tri_ids;
    /* informational (FastNodeManager::averageBvhDepth / triangleCount) */
    float average_depth;
    uint32_t triangle_count;
    uint32_t max_depth;
} rt_bvh_desc;

typedef struct rt_scene* rt_scene_t;

/* ---- trace flags ------------------------------------------------------------ */
#define RT_CLOSEST 0u      /* FastNodeManager::intersect            (:1328-1418) */
#define RT_ANY 1u          /* FastNodeManager::intersectSecondary   (:1420-1493) */
#define RT_SAVE_DIST 2u    /* FastNodeManager::intersectSaveDistance(:1232-1326) */
#define RT_STREAM 4u       /* ray-stream scheduling (intersectWideAlternative
                              :835-1049), per-ray results identical to the
                              per-ray traversal                                   */
#define RT_COUNTS 8u       /* write rt_counts per ray                             */
#define RT_DEVICE_PTRS 16u /* rays/out/counts are device pointers                 */
#define RT_NO_SYNC 32u     /* with RT_DEVICE_PTRS: return without waiting         */
#define RT_COOP 64u        /* wavefront-cooperative child tests (lanes = slots)   */
#define RT_EXACT_BOXES 256u /* per-ray closest / any hit: test every child on its
                               exact box at the parent (the exact-box node
                               records), not on the quantized records; results
                               are the same either way                            */

/* ---- scene lifetime ----------------------------------------------------------- */

/* Converts the FastNode arrays into the device layout and copies them into
 * HBM on the current device.  The desc arrays are only read during the call. */
int rt_upload_bvh(const rt_bvh_desc* desc, rt_scene_t* out);

/* Traces n rays.  out is rt_hit[n] for RT_CLOSEST / RT_SAVE_DIST and
 * uint8_t[n] (1 = occluded) for RT_ANY.  counts may be NULL unless RT_COUNTS.
 * hip_stream is a hipStream_t.  NULL keeps the null stream's order: the
 * work starts after what is already queued on the null stream, and with
 * RT_NO_SYNC later null-stream work waits for it (it runs on a stream of
 * the library's).  With
 * RT_DEVICE_PTRS the buffers stay on the device; otherwise they are staged.
 * n == 0 returns RT_OK without reading any buffer.  Flag bits other than the
 * RT_* trace flags above, RT_ANY with RT_SAVE_DIST, and RT_COUNTS without a
 * counts buffer return RT_ERR_ARG; RT_SAVE_DIST with RT_STREAM or RT_COOP,
 * and RT_COOP with RT_STREAM, return RT_ERR_UNSUPPORTED.
 * Any number of host threads may call into one scene at once (the ISPC
 * kernels this replaces are called from OpenMP workers, cameraFast.h:320,465):
 * the scene serialises the enqueueing of the launches, a call that waits for
 * its results waits without the scene's lock (so calls on different streams
 * overlap on the device), each hip_stream gets its own scratch, and
 * rt_last_error() reports the calling thread's own error.  A scene keeps the
 * scratch of 16 streams; when another stream arrives, the least recently
 * used scratch that no call is waiting on passes to it after its last launch
 * (when every kept scratch is waited on, the set grows until those calls
 * return). */
int rt_trace(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags,
             void* out, rt_counts* counts, void* hip_stream);

/* rt_trace for callers that cannot allocate: code under HIP graph capture,
 * serving loops and harnesses that forbid hipMalloc.  No counterpart in the
 * reference (its ISPC kernels keep their state on the caller's stack).
 * The per-ray kernel's scratch (work counters and stack spill) lives in one
 * fixed array of RT_NOALLOC_SCRATCH_BYTES per GPU, held from the library's
 * load until the process ends; the call allocates nothing, records no event
 * and does not wait: it enqueues a reset kernel and the trace kernel on
 * hip_stream (NULL = the null stream) and returns.  rays, out and counts are
 * device pointers (RT_DEVICE_PTRS and RT_NO_SYNC are implied); flags:
 * RT_CLOSEST / RT_ANY / RT_SAVE_DIST and RT_COUNTS; RT_STREAM and RT_COOP
 * return RT_ERR_UNSUPPORTED.  Results are those of rt_trace.  The launch
 * uses as many persistent blocks as rt_trace up to what the spill fits; a
 * scene whose stack spill for one block exceeds the array returns
 * RT_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array:
 * calls whose kernels overlap on one GPU (different streams, or host
 * threads) race and must be ordered by the caller.  The first call on each
 * GPU resolves the array's device address (hipGetSymbolAddress): make it
 * outside graph capture.  Rays stopped by a guard are not added to
 * rt_stats.error_rays (nothing is read back). */
#define RT_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int rt_trace_noalloc(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                     rt_counts* counts, void* hip_stream);

typedef struct rt_stats_t {
    uint64_t bytes_resident;  /* device bytes owned by the scene               */
    uint32_t inner_nodes, leaves;
    uint32_t node_record_bytes, leaf_block_bytes;
    float last_kernel_ms;     /* duration of the last trace kernel (HIP events) */
    uint64_t rays_traced;     /* rays traced through this scene so far          */
    uint32_t block_size;      /* launch configuration chosen for the scene     */
    uint32_t stack_window;    /* LDS stack entries per lane                    */
    uint32_t max_stack;       /* stack capacity per ray (LDS + spill)          */
    uint32_t error_rays;      /* rays stopped by a guard: stack capacity exceeded,
                                 or a padded child id past the node array; the
                                 counters accumulate on the device and rt_stats
                                 waits for launches still in flight to read them */
    uint32_t quantized_record_bytes; /* inner record of the quantized node format
                                 the per-ray closest / any-hit kernels use; 0 when
                                 the scene does not qualify (box coordinates
                                 beyond +-2^60, a child box outside its parent's,
                                 padded slots that are not zero boxes) and every
                                 kernel tests the exact boxes                     */
} rt_stats_t;

int rt_stats(rt_scene_t s, rt_stats_t* out);

/* Visit statistics of one traced batch (the counters of the reference's
 * instrumented renderer, cameraData.h:241-329, kept over this traversal):
 * visits of inner nodes and leaves by depth (root = depth 0), visits by the
 * visited node's child count and the visited leaf's triangle count, child
 * boxes tested / hit, and triangles tested / passing the test.  Tests count
 * real children and triangles (not the padded slots).  Totals, not per ray. */
#define RT_HIST_DEPTHS 64  /* depths >= 63 share the last bin */
#define RT_HIST_FULL 17
typedef struct rt_trace_hist {
    uint64_t rays;
    uint64_t node_visits, leaf_visits;
    uint64_t aabb_tests, aabb_hits;
    uint64_t tri_tests, tri_hits;
    uint64_t node_depth[RT_HIST_DEPTHS];
    uint64_t leaf_depth[RT_HIST_DEPTHS];
    uint64_t child_fullness[RT_HIST_FULL];
    uint64_t leaf_fullness[RT_HIST_FULL];
} rt_trace_hist;

/* Traces like rt_trace (flags: RT_CLOSEST / RT_ANY / RT_SAVE_DIST,
 * RT_DEVICE_PTRS) with the instrumented kernel and fills *hist (host memory;
 * the call waits for the result).  `out` receives the same hits / occlusion
 * flags as rt_trace. */
int rt_trace_histograms(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                        rt_trace_hist* hist, void* hip_stream);

/* Per-step workload of ray tiles: replaces the instrumented wide traversal
 * CompactNodeManager::intersectWideAlternative (compactNode.cpp:358-601),
 * whose per-workgroup vectors cameraData.h:660-1109 writes out.  The rays
 * form tiles of `tile` consecutive rays (a camera workgroup of wg x wg
 * pixels in CameraFast tile order); n must be a multiple of `tile`.
 * Within a tile every unfinished ray takes exactly one visit per step (one
 * node or one leaf, in its own depth-first order, as rt_trace with RT_CLOSEST
 * or RT_ANY traverses it), and each step records how many rays visit a node
 * or a leaf, how many distinct nodes and leaves they visit, and how many rays
 * finish.  Record 0 is the state before the first step: node_work = rays
 * started, node_unique = 1, leaf_* = 0, terminations = rays not started.
 * With RT_ANY a ray whose origin x is NaN (a pixel without a primary hit)
 * is not started, as in the reference; with RT_CLOSEST every ray starts.
 * out: (n / tile) x step_cap records; steps[i] = records of tile i (its
 * longest ray's visits + 1).  A tile with more than step_cap records
 * writes the first step_cap and reports its full count in steps[i]: call
 * again with step_cap = max(steps).  counts (optional) receives each ray's
 * visits (zero for a ray not started).  flags: RT_CLOSEST or RT_ANY, plus
 * RT_DEVICE_PTRS; the call waits for the result.  tile <= 1024. */
typedef struct rt_step_stat {
    uint32_t node_work, node_unique, leaf_work, leaf_unique, terminations;
} rt_step_stat;
int rt_trace_steps(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, uint32_t tile,
                   uint32_t step_cap, rt_step_stat* out, uint32_t* steps, rt_counts* counts, void* hip_stream);

/* Self-test of the one arithmetic primitive the kernels do not take from the
 * compiler: the reciprocal of the leaf test (inv = 1.f / div,
 * rayTracerTemplate.ispc:59, correctly rounded like the ISPC code's divps).
 * The kernels form it from the hardware reciprocal and one Newton step; this
 * compares it with the correctly rounded division 1.0f / x for the `count`
 * float bit patterns first_bits, first_bits + 1, ... (wrapping at 2^32,
 * count <= 2^32) on the current device.  *mismatches = patterns whose
 * results differ in any bit (two NaNs count as equal); *first_mismatch
 * (optional) = the lowest such pattern, or UINT64_MAX.  values (optional,
 * device, count floats) receives the reciprocals.  No reference
 * counterpart; the call waits for the result. */
int rt_check_recip(uint32_t first_bits, uint64_t count, float* values, uint64_t* mismatches,
                   uint64_t* first_mismatch, void* hip_stream);

/* Load-pipeline calibration on the current device: wave-level 16-byte loads
 * per second when each of the first active_lanes (1..64) lanes of every wave
 * loads from its own 128-byte line of a buffer_bytes buffer (a power of two
 * in [8192, 2^30]; 16 KB stays in the vector L1), 6 waves per SIMD.  This is
 * the shape of the traversal kernels' node and leaf loads, one record per
 * lane, and the rate is the ceiling their load count is measured against.
 * No reference counterpart; the call allocates, runs a warm-up launch and
 * three timed ones (best of three, about 0.15 s in all) on a stream of its
 * own and waits. */
int rt_measure_gather(uint32_t active_lanes, uint32_t buffer_bytes, double* wave_loads_per_s);

/* Overrides the launch configuration (block size 64, 128, 192 or 256,
 * LDS stack window 8/16/32 entries).  0 keeps the current value. */
int rt_set_launch(rt_scene_t s, uint32_t block_size, uint32_t stack_window);

/* Node format of this scene's items in batched launches with
 * RT_SCENE_SHAPES: 0 = the quantized records when the scene has them (the
 * default), RT_EXACT_BOXES = the exact-box records.  Single-scene calls
 * choose theirs per call (RT_EXACT_BOXES).  Results are the same either way;
 * the autotuner sets the faster one.  No reference counterpart. */
int rt_set_node_format(rt_scene_t s, uint32_t flags);

/* Wave schedule of the per-ray kernel.  0: every lane takes one step per
 * iteration, so a wave runs the node and the leaf code whenever its lanes
 * disagree.  r > 0: each iteration runs one phase only, the leaf phase when
 * 16 * (lanes at a leaf) >= r * (lanes at a node); the other lanes wait.
 * Every ray still takes its own steps in its own order, so results and
 * counts do not depend on r.  No counterpart in the reference (its
 * traversal is one ray per call); tuned per scene like the launch shape. */
int rt_set_schedule(rt_scene_t s, uint32_t leaf_ratio16);

int rt_release(rt_scene_t s);

const char* rt_last_error(void);

/* Number of device bytes a scene built from desc would occupy (admission). */
int rt_scene_footprint(const rt_bvh_desc* desc, uint64_t* bytes);

/* ---- batched launch ------------------------------------------------------------------ */

/* Many scenes and ray batches of uneven size in ONE persistent launch: the
 * rays of all items form one index space that the waves drain through a
 * shared work counter (work stealing at ray granularity); every lane carries
 * the scene of its current ray.  Scenes of different N may share a batch:
 * the items are grouped by node memory, one counter per group, and each
 * group runs code specialised on its node memory.  No reference counterpart: the reference
 * traces one scene per CameraFast render (cameraFast.h:318-582). */
typedef struct rt_batch* rt_batch_t;

typedef struct rt_batch_item {
    rt_scene_t scene;
    const rt_ray* rays; /* device pointer, n rays                              */
    void* out;          /* device: rt_hit[n] (RT_CLOSEST) or uint8_t[n] (RT_ANY) */
    rt_counts* counts;  /* device, optional (NULL)                              */
    uint32_t n;
    uint32_t flags;     /* RT_CLOSEST or RT_ANY                                 */
} rt_batch_item;

#define RT_GRAPH 128u /* rt_batch_trace: replay the captured hipGraph (re-armed
                         with the new item table) instead of a direct launch  */
#define RT_SCENE_SHAPES 512u /* rt_batch_trace: every scene's items run with that
                         scene's own launch shape (rt_set_launch), wave schedule
                         (rt_set_schedule) and node format (rt_set_node_format);
                         the items are grouped into one persistent launch per
                         distinct (format, block size, window, schedule), at
                         most 16, which run concurrently (independent branches
                         of one graph, or forked streams joined again)        */

int rt_batch_create(uint32_t max_items, rt_batch_t* out);
/* Items must all live on the batch's device; total rays < 2^31.  A NULL
 * hip_stream keeps the null stream's order, as for rt_trace.  flags takes
 * RT_GRAPH and RT_NO_SYNC (other bits: RT_ERR_ARG); items with n == 0 are
 * skipped.  The kernel stages the batch's tables in LDS beside its stack
 * window (48 bytes per distinct scene, 16 per item, rounded up to 256): a
 * batch whose tables do not fit with the launch shape returns RT_ERR_ARG
 * (thousands of items fit; larger tables lower the kernel's occupancy). */
int rt_batch_trace(rt_batch_t b, const rt_batch_item* items, uint32_t n_items, uint32_t flags,
                   void* hip_stream);
int rt_batch_set_launch(rt_batch_t b, uint32_t block_size, uint32_t stack_window);
int rt_batch_stats(rt_batch_t b, rt_stats_t* out);
int rt_batch_release(rt_batch_t b);

/* ---- serving ----------------------------------------------------------------------- */

/* A request loop over resident scenes with continuous batching: requests are
 * admitted first-come first-served while the device bytes of admitted
 * requests (rays + results) stay under hbm_cap_bytes, and everything admitted
 * and waiting is coalesced into one batched persistent launch whenever one of
 * `slots` launch slots (stream + device arena) is free, so requests that arrive
 * while others run join the next launch and overlap with the running ones.
 * Latency is measured from rt_server_submit to the results being in the
 * caller's buffer.  No reference counterpart (the reference renders offline). */
typedef struct rt_server* rt_server_t;

typedef struct rt_server_config {
    uint64_t hbm_cap_bytes; /* admission cap on admitted request bytes (0 = 4 GiB) */
    uint32_t slots;         /* concurrent launches (0 = 2)                         */
    uint32_t max_items;     /* requests per launch (0 = 512)                       */
    uint32_t block_size;    /* launch configuration (0 = defaults)                 */
    uint32_t stack_window;
    uint32_t use_graph;     /* replay captured hipGraphs (1) or launch directly (0)  */
    uint32_t copy_threads;  /* host threads that copy request rays into the pinned
                               staging pool and results out of it (0 = 8)       */
} rt_server_config;

typedef struct rt_server_stats {
    uint64_t submitted, completed, failed, rejected, cancelled;
    uint64_t rays_completed, launches;
    double p50_ms, p99_ms, mean_ms, max_ms; /* request latency since the last reset */
    double wall_s;        /* first submit to last completion since the last reset  */
    double mrays_per_s;   /* rays_completed / wall_s                                */
    uint64_t peak_admitted_bytes;
    double mean_requests_per_launch;
} rt_server_stats;

int rt_server_create(const rt_server_config* cfg, rt_server_t* out);
/* Registers a resident scene (the server does not own it). */
int rt_server_add_scene(rt_server_t sv, rt_scene_t scene, uint32_t* scene_id);
/* Queues a request.  rays/out are host buffers that must stay valid until the
 * request completes (out: rt_hit[n] for RT_CLOSEST, uint8_t[n] for RT_ANY).
 * RT_ERR_BUSY: the request alone exceeds the admission cap.  With
 * RT_DEVICE_PTRS, rays/out are device buffers on the server's device (rays
 * made by rt_spawn_*_device, results consumed on the device): the launch reads
 * and writes them in place, with no host copies and no bytes of the pool (the
 * admission cap bounds host requests only); they are admitted in arrival
 * order like the others, and rt_server_wait returns once `out` is written. */
int rt_server_submit(rt_server_t sv, uint32_t scene_id, const rt_ray* rays, uint32_t n,
                     uint32_t flags, void* out, uint64_t* request_id);
/* RT_OK when the request completed, RT_ERR_BUSY on timeout, the launch's error
 * code if its launch failed, RT_ERR_ARG for an unknown/cancelled id. */
int rt_server_wait(rt_server_t sv, uint64_t request_id, double timeout_s);
/* Cancels a request that has not been launched yet. */
int rt_server_cancel(rt_server_t sv, uint64_t request_id);
int rt_server_drain(rt_server_t sv);
int rt_server_get_stats(rt_server_t sv, rt_server_stats* out, int reset);
/* Stops the dispatcher, waits for launched work and drops queued requests
 * (their output buffers are not written).  No other thread may be inside an
 * rt_server_* call on this server when it is called. */
int rt_server_destroy(rt_server_t sv);

/* ---- host side: scenes, builder, camera, config ------------------------------- */

/* Procedural scenes (PCG32 seeded).  kind 1: spheres + ground plane
 * (icospheres of frequency `param`, `count` of them); kind 2: clustered
 * "interior" room (count = total triangles, param = clusters, 0 = auto).
 * Triangles are returned as float[n*9] (p0,p1,p2) owned by the library
 * (free with rt_free). */
#define RT_SCENE_SPHERES 1u
#define RT_SCENE_INTERIOR 2u
int rt_scene_generate(uint32_t kind, uint64_t seed, uint32_t count, uint32_t param,
                      float** tris, uint32_t* n_tris);
void rt_free(void* p);

/* Bvh(prims, N, L, sortEachSplit, leafSplitOption) + recursiveBvh(sahFactor)
 * (bvh.cpp:28-40, aabb.cpp:71-206) followed by the FastNodeManager flattener
 * (fastNodeManager.h:1495-1688), with primitives pre-sorted as
 * RayTracer::preparePrimitives does (rayTracer.cpp:747-754). */
typedef struct rt_build_params {
    uint32_t branching, leaf_size, gang;
    uint32_t sort_each_split;   /* config.txt sortEachSplit (1)   */
    int32_t leaf_split_option;  /* config.txt leafSplitOption (1) */
    float sah_factor;           /* loadSahFactor(L, N, G)         */
    uint32_t threads;           /* 0 = all cores                  */
    uint32_t node_memory;       /* 0 = roundup(N, G); else a larger multiple of G,
                                   the node memory test (rayTracer.cpp:41-52)   */
    uint32_t leaf_memory;       /* 0 = roundup(L, G); else a larger multiple of G,
                                   the leaf memory test (rayTracer.cpp:630-640) */
} rt_build_params;

typedef struct rt_host_bvh* rt_host_bvh_t;
int rt_build(const float* tris, uint32_t n_tris, const rt_build_params* p, rt_host_bvh_t* out);
/* Fills desc with views into h (valid until rt_host_bvh_free). */
int rt_host_bvh_desc(rt_host_bvh_t h, rt_bvh_desc* desc);
int rt_host_bvh_free(rt_host_bvh_t h);

/* Pinhole camera (camera.h:52-58,91-98): rays in CameraFast tile order
 * (tile i, lane j -> ray i*wg*wg + j, cameraFast.h:324-342); rows of a
 * partial last tile row are skipped.  Writes width*height rays. */
typedef struct rt_camera {
    float pos[3], target[3], up[3];
    float focal; /* 0.866 = 60 degree horizontal FOV */
} rt_camera;
int rt_camera_rays(const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                   rt_ray* out);

/* Hit reconstruction (getSurfaceNormalPosition, fastNodeManager.h:1699-1726):
 * for each hit writes normal[3] and position[3]; misses get NaN. */
int rt_hit_surface(rt_host_bvh_t h, const rt_ray* rays, const rt_hit* hits, uint32_t n,
                   float* normal_out, float* position_out);

/* Shadow rays toward a point light: FastRay(P + n*0.001, light, tmax).
 * Returns the number of rays written (one per hit); src_index[i] = ray id. */
int rt_spawn_shadow(const float* normal, const float* position, const rt_hit* hits,
                    uint32_t n, const float light[3], float tmax, rt_ray* out,
                    uint32_t* src_index, uint32_t* n_out);

/* Ambient-occlusion rays of one sample (the performance renderer's ambient
 * loop, cameraFast.h:395-418, Camera::getAmbientDirection, camera.cpp:21-52):
 * for every hit pixel i, in pixel order, FastRay from P + n*0.001 with a
 * hashed direction uniform on the sphere, flipped into the normal's
 * hemisphere (the hash is this library's, not std::hash), tMax = dist.
 * Writes *n_out rays; src_index[k] = pixel of ray k (optional). */
int rt_spawn_ambient(const float* normal, const float* position, const rt_hit* hits, uint32_t n,
                     uint32_t sample, float dist, rt_ray* out, uint32_t* src_index, uint32_t* n_out);

/* n_out cosine-distributed bounce rays cycling over the hits (PCG32 stream
 * per ray index); the normal is face-forwarded against the primary ray. */
int rt_spawn_diffuse(const rt_ray* primary, const float* normal, const float* position,
                     const rt_hit* hits, uint32_t n, uint64_t seed, uint32_t n_out,
                     float tmax, rt_ray* out);

/* ---- Ray generation and hit shading on the GPU ----------------------------
 * Device pointers only; each call enqueues its kernels on hip_stream (NULL =
 * the default stream) and returns.  The arithmetic is shared with the host
 * functions above, so the results are bit-identical to theirs. */

/* Primary rays in CameraFast tile order, as rt_camera_rays
 * (Camera::getRayTargetPosition, camera.h:91-98; cameraFast.h:318-342).
 * out: width*height rays. */
int rt_camera_rays_device(const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                          rt_ray* out, void* hip_stream);

/* Hit reconstruction from the scene's resident triangles, as rt_hit_surface
 * (getSurfaceNormalPosition, fastNodeManager.h:1699-1726); misses get NaN. */
int rt_hit_surface_device(rt_scene_t s, const rt_ray* rays, const rt_hit* hits, uint32_t n,
                          float* normal_out, float* position_out, void* hip_stream);

/* Shadow rays toward a point light, one per input: ray i = rt_spawn_shadow's
 * ray for hit i (FastRay(P + n*0.001, light, tmax), cameraFast.h:373,521).
 * Where hit i is a miss, ray i has tmax = -1, which no box test accepts: it
 * costs one root visit and is never occluded.  (The reference gives such
 * pixels a NaN origin, cameraFast.h:522-527, and skips them in its stream
 * any-hit; the per-ray traversal would accept every box for a NaN origin.) */
int rt_spawn_shadow_device(const float* normal, const float* position, const rt_hit* hits, uint32_t n,
                           const float light[3], float tmax, rt_ray* out, void* hip_stream);

/* n_out bounce rays exactly as rt_spawn_diffuse (ray r from hit number
 * r % hits, in pixel order); with no hit at all every ray gets tmax = -1
 * (no hit, as above).
 * Allocates its scratch stream-ordered (hipMallocAsync) on hip_stream. */
int rt_spawn_diffuse_device(const rt_ray* primary, const float* normal, const float* position,
                            const rt_hit* hits, uint32_t n, uint64_t seed, uint32_t n_out, float tmax,
                            rt_ray* out, void* hip_stream);

/* One frame of the performance renderer on the GPU (CameraFast::renderImage,
 * cameraFast.h:318-461): the primary rays of cam in tile order traced
 * (flags: RT_CLOSEST or RT_SAVE_DIST, plus RT_STREAM), their hit surfaces,
 * `samples` ambient rays per hit pixel (rt_spawn_ambient's rays, any hit,
 * with RT_STREAM if given), and per pixel the gray value (uint8)((1 -
 * occluded / samples) * 255), 0 for a miss, written to gray_out[width *
 * height] (host memory, tile order).  ms_out (optional, host, 2 floats) =
 * {primary rays, whole frame} in ms by HIP events.  The result equals the
 * host pipeline's (rt_camera_rays, rt_trace, rt_hit_surface,
 * rt_spawn_ambient).  Device scratch is allocated stream-ordered
 * (hipMallocAsync); the call waits for the frame. */
int rt_render_ambient(rt_scene_t s, const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                      uint32_t flags, uint32_t samples, float dist, uint8_t* gray_out, float* ms_out,
                      void* hip_stream);

/* sahFactorConfigSse.txt lookup (RayTracer::loadSahFactor, rayTracer.cpp:880-925). */
float rt_sah_factor(uint32_t leaf_size, uint32_t branching, uint32_t gang, const char* table_path);

#ifdef __cplusplus
}
#endif
#endif /* BVHRT_RT_H */
