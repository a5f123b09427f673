// Device hit reconstruction (rt_shade.hip), launched by rt_hit_surface_device
// with the scene's device arrays.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rt.h"

namespace rt {

// tris: device leaf blocks (rt_layout.h); leafLc: loop count per leaf block
// as getLeafLoopCount reads it (0 = unusable).
hipError_t launch_surface(const float* tris, const uint8_t* leafLc, uint32_t Lm, uint32_t nLeaves,
                          const rt_ray* rays, const rt_hit* hits, uint32_t n, float* nrm, float* pos,
                          hipStream_t stream);

// The stream rt_trace uses for scene s when the caller passes none (created
// on first use, ordered after the null stream's queued work); calls that
// enqueue several steps use it too, so every step runs on one stream.  Makes
// the scene's device current.
int scene_stream(rt_scene_t s, hipStream_t* out);

// Device a scene lives on (-1 for a null scene).
int scene_device(rt_scene_t s);

// The per-ray trace (flags: RT_CLOSEST / RT_ANY / RT_SAVE_DIST, RT_COUNTS;
// device pointers) on stream st with its work counters and stack spill in
// the caller's scratchBytes of device memory at scratchBase (256-byte
// aligned): no allocation, no event, no wait.  The grid is trimmed so the
// spill fits; RT_ERR_UNSUPPORTED when one block's spill does not.  `who`
// names the entry point in error messages.
int trace_in_scratch(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out, rt_counts* counts,
                     hipStream_t st, uint32_t* scratchBase, size_t scratchBytes, const char* who);

}  // namespace rt
