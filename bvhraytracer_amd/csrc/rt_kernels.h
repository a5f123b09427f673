// Traversal kernels: per-ray wide-BVH traversal on CDNA4, one ray per lane,
// persistent waves refilling lanes from a global ray counter.
//
// Semantics follow the reference traversals exactly (bit-exact hit identity
// and node/leaf counts), with the FastNode arrays already converted to the
// device layout of rt_layout.h:
//   closest hit  FastNodeManager::intersect            fastNodeManager.h:1328-1418
//   save dist    FastNodeManager::intersectSaveDistance fastNodeManager.h:1232-1326
//   any hit      FastNodeManager::intersectSecondary    fastNodeManager.h:1420-1493
// with the ISPC kernels inlined:
//   slab test    aabbIntersect{k}   rayTracerTemplate.ispc:170-234
//   closest MT   triIntersect{k}    rayTracerTemplate.ispc:25-108
//   any-hit MT   triAnyHit{k}       rayTracerTemplate.ispc:111-168
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rt.h"
#include "rt_layout.h"

namespace rt {

enum TraceMode { kClosest = 0, kAny = 1, kSaveDist = 2, kBatch = 3 };

// Work counters: the ray counter and the guarded-condition counter, each on
// its own 128-byte line.  launch_reset zeroes the ray counter's line before
// every launch; the guard counter accumulates from the block's creation
// (zeroed then with launch_zero) and is read by the statistics calls.
constexpr uint32_t kRayCounter = 0;
constexpr uint32_t kErrWord = 32;  // rays stopped by a guard
constexpr uint32_t kWorkWords = 64;
// Batched launch grouped by node memory and mode (closest, any hit): group
// g's ray counter is word kGroupCounter + kGroupCounterStride * g (inside the
// line launch_reset zeroes; group 0 shares the plain counter's word).
constexpr uint32_t kMaxGroups = 8;
constexpr uint32_t kGroupCounter = kRayCounter;
constexpr uint32_t kGroupCounterStride = 4;
static_assert(kGroupCounter + kGroupCounterStride * (kMaxGroups - 1) < kErrWord, "group counters overlap");

// Group key: node memory and mode; ascending keys order the groups.
constexpr uint32_t batch_group_key(uint32_t M, uint32_t mode) { return M << 1 | (mode == kAny ? 1u : 0u); }

// Batched launch tables (device memory).  A batch is a list of items, each a
// contiguous ray range of one scene in one mode; the global ray index space
// is the concatenation of the items (item i covers [start, start + n)).
struct BatchScene {
    const uint32_t* nodes;
    const float* tris;
    const int32_t* triIds;
    const uint32_t* nodeEntry;
    uint32_t rootEntry, gang, Lm, nodeCount, M, boxesBounded, pad1, pad2;
};

struct BatchItem {
    const rt_ray* rays;
    void* out;          // rt_hit[n] (closest) or uint8_t[n] (any)
    rt_counts* counts;  // optional
    uint32_t scene, mode, start, n;
};

struct KParams {
    const uint32_t* __restrict__ nodes;  // inner records
    const float* __restrict__ tris;      // leaf blocks
    const int32_t* __restrict__ triIds;  // leaf ordinal * Lm + slot -> triangle id
    const uint32_t* __restrict__ nodeEntry;  // FastNode id -> stack entry (rare path)
    const rt_ray* __restrict__ rays;
    rt_hit* hits;
    uint8_t* occluded;
    rt_counts* counts;
    uint32_t* spill;       // per thread slot: spillCap entries (+ spillCap floats)
    uint32_t* work;        // kWorkWords words: partition counters, guard counter
    uint32_t n;
    uint32_t rootEntry;
    uint32_t spillCap;
    uint32_t gang;
    uint32_t Lm;
    uint32_t nodeCount;
    uint32_t chunk;  // rays reserved per wave per atomicAdd on the counter
    uint32_t boxesBounded;  // every tested box coordinate is within +-2^100
    // visit statistics launch only: real triangles per leaf block, and the
    // rt_trace_hist counters (uint64) the kernel adds to
    const uint8_t* leafPrims;
    uint64_t* hist;
    // per-ray kernel wave schedule: 0 = every lane steps each iteration;
    // r > 0 = one phase per iteration, the leaf phase when
    // 16 * leafLanes >= r * nodeLanes
    uint32_t sched;
    // cooperative launch only: node memory and LDS stack entries per ray
    uint32_t M;
    uint32_t coopDepth;
    // per-step tile statistics launch only: stepCap records per tile, and
    // the record count of each tile
    rt_step_stat* steps;
    uint32_t* stepLen;
    uint32_t stepCap;
    // batched launch only
    const BatchItem* items;
    const BatchScene* scenes;
    uint32_t nItems;
    uint32_t nScenes;
    // batched launch: items ordered by node memory and mode into nGroups
    // groups, group g holding global rays [groupEnd[g-1], groupEnd[g]) of
    // group key groupKey[g] (batch_group_key)
    uint32_t nGroups;
    uint32_t groupKey[kMaxGroups];
    uint32_t groupEnd[kMaxGroups];
};

// Node memory argument of a batched launch over scenes of several node
// memories (the grouped kernel): all of them up to 16, or all up to 8.
constexpr uint32_t kBatchMixed = 0;
constexpr uint32_t kBatchMixedSmall = 1;

// Launches the per-ray kernel for node memory M on `stream`; mode kBatch
// launches the batched kernel (M = kBatchMixed or kBatchMixedSmall: the
// grouped one).
// quantized: p.nodes (every scene's nodes, batched) are the quantized
// records and p.spill has a second region of spillCap floats per slot
// (closest, any and batched launches).
// Batched launches (mode kBatch) stage their scene and item tables in LDS:
// tableBytes >= batch_table_bytes(p.nScenes, p.nItems) more dynamic LDS.
hipError_t launch_trace(uint32_t M, int mode, uint32_t stackWindow, uint32_t blockSize,
                        uint32_t grid, const KParams& p, hipStream_t stream, bool quantized = false,
                        size_t tableBytes = 0);

// Blocks per CU the per-ray kernel reaches with this configuration.
int trace_occupancy(uint32_t M, int mode, uint32_t stackWindow, uint32_t blockSize, bool quantized = false,
                    size_t tableBytes = 0);

// LDS bytes of a batched launch's staged tables.
size_t batch_table_bytes(uint32_t nScenes, uint32_t nItems);

// Kernel functions for graphs built node by node: the batched kernel for
// node memory M (0: the grouped one), window S and node format (grid and
// block as launch_trace, dynamic LDS trace_lds_bytes(M, kBatch, ...) plus the
// tables; one KParams argument), and the reset kernel (1 block of 64
// threads, one uint32_t* argument: the work words).
const void* batch_kernel(uint32_t M, uint32_t S, bool quantized);
const void* reset_kernel();

// Dynamic LDS bytes of the per-ray kernel (stack window, and for the
// save-distance mode the stack distances and aabbDistances, for the
// quantized records the entries' tmax).
size_t trace_lds_bytes(uint32_t M, int mode, uint32_t stackWindow, uint32_t blockSize, bool quantized = false);

// Ray-stream kernel (closest / any): blocks of stream_block() ray slots.
hipError_t launch_stream(uint32_t M, int mode, uint32_t stackWindow, uint32_t grid, const KParams& p,
                         hipStream_t stream);
int stream_occupancy(uint32_t M, int mode, uint32_t stackWindow);
uint32_t stream_block();

// Cooperative kernel (closest / any): 256-thread blocks, W lanes per ray.
uint32_t coop_width(uint32_t M, uint32_t Lm);
hipError_t launch_coop(uint32_t W, int mode, uint32_t grid, const KParams& p, hipStream_t stream);
int coop_occupancy(uint32_t W, int mode, uint32_t depth);

// Zeroes the ray counter before a launch (not the guard counter).
hipError_t launch_reset(uint32_t* work, hipStream_t stream);

// Self-test of the leaf test's reciprocal against the correctly rounded
// division over `count` float bit patterns from `first` (wrapping):
// counter[0] += mismatches, counter[1] = min(counter[1], lowest mismatching
// pattern); values (optional) receives the reciprocals.
hipError_t launch_check_recip(uint32_t first, uint64_t count, float* values, unsigned long long* counter,
                              hipStream_t stream);
// Load-pipeline calibration: `blocks` blocks of 256 threads; in every wave
// the first `active` lanes each load 16 bytes from their own 128-byte line of
// buf (lines * 128 bytes, lines a power of two >= 64, zero-filled), `iters`
// times; sink: 64 words.
hipError_t launch_gather_probe(const void* buf, uint32_t lines, uint32_t active, uint32_t iters, uint32_t blocks,
                               uint32_t* sink, hipStream_t stream);
// Zeroes `words` 32-bit words with a kernel (no memset node if captured).
hipError_t launch_zero(uint32_t* p, uint32_t words, hipStream_t stream);

// Instrumented per-ray kernel (closest / any / save distance): blocks of
// trace_hist_block() threads with a 16-entry LDS stack window; p.hist and
// p.leafPrims set.
constexpr uint32_t kHistWindow = 16;
hipError_t launch_trace_hist(uint32_t M, int mode, uint32_t grid, const KParams& p, hipStream_t stream);
uint32_t trace_hist_block(uint32_t M, int mode);
int trace_hist_occupancy(uint32_t M, int mode);

// Per-step tile statistics kernel (closest / any): blocks of `tile` threads,
// one tile per block iteration, an 8-entry LDS stack window; p.steps,
// p.stepLen and p.stepCap set, p.n a multiple of tile.
constexpr uint32_t kStepWindow = 8;
hipError_t launch_trace_steps(uint32_t M, int mode, uint32_t tile, uint32_t grid, const KParams& p,
                              hipStream_t stream);
int trace_steps_occupancy(uint32_t M, int mode, uint32_t tile);

}  // namespace rt
