// rt_trace_noalloc: rt_trace for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops that forbid hipMalloc).  The
// per-ray kernel's scratch (work counters, stack spill) lives in a fixed
// device array instead of the per-stream buffers rt_trace allocates.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "host_common.h"
#include "rt_kernels.h"
#include "rt_shade.h"

namespace {

constexpr size_t kScratchBytes = size_t(RT_NOALLOC_SCRATCH_BYTES);
static_assert(kScratchBytes % 256 == 0, "scratch size");

// Work counters, then the stack spill.  Every GPU that runs the library
// holds its own copy until the process ends.
__device__ __align__(256) uint32_t scratch[kScratchBytes / 4];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The address of `scratch` on device dev, resolved with dev current on the
// first call for that device.  In host code the array's name is a host-side
// stand-in, never a device address: only this pointer reaches kernels.
int scratch_ptr(int dev, uint32_t** out) {
    if (dev < 0 || dev >= kMaxDevices) return rt::set_error(RT_ERR_UNSUPPORTED, "rt_trace_noalloc: device index above 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) (void)hipSetDevice(prev);
        if (e != hipSuccess)
            return rt::set_error(RT_ERR_HIP, std::string("rt_trace_noalloc: hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(e));
        if (!p) return rt::set_error(RT_ERR_HIP, "rt_trace_noalloc: no device address for the scratch array");
        g_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_ptr[dev];
    return RT_OK;
}

}  // namespace

extern "C" int rt_trace_noalloc(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                                rt_counts* counts, void* hip_stream) {
    if (!s) return rt::set_error(RT_ERR_ARG, "rt_trace_noalloc: null scene");
    uint32_t* base = nullptr;
    const int rc = scratch_ptr(rt::scene_device(s), &base);
    if (rc != RT_OK) return rc;
    return rt::trace_in_scratch(s, rays, n, flags, out, counts, static_cast<hipStream_t>(hip_stream), base,
                                kScratchBytes, "rt_trace_noalloc");
}
