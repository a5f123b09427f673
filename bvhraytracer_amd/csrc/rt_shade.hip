// Ray generation and hit shading on the GPU: primary rays in CameraFast tile
// order, hit reconstruction, shadow and diffuse secondary rays.  The
// arithmetic is rt_raymath.h, shared with the host entry points, so the
// device results are bit-identical to rt_camera_rays / rt_hit_surface /
// rt_spawn_shadow / rt_spawn_diffuse.  All of these kernels are elementwise
// and bound by HBM bandwidth (40-64 bytes per ray).
#include <hipcub/hipcub.hpp>

#include "host_common.h"
#include "rt_kernels.h"
#include "rt_layout.h"
#include "rt_raymath.h"
#include "rt_shade.h"

namespace rt {
namespace {

constexpr uint32_t kShadeBlock = 256;

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kShadeBlock - 1) / kShadeBlock); }

// One thread per tile lane; pixels below the image are skipped and the
// output index is the lane's position among the kept pixels (only the last
// tile row is partial, so that is tile * wg^2 until the last row).
__global__ void __launch_bounds__(kShadeBlock) k_camera(CameraBasis b, uint64_t lanes, rt_ray* out) {
    const uint64_t g = (uint64_t)blockIdx.x * kShadeBlock + threadIdx.x;
    if (g >= lanes) return;
    const uint32_t wgsq = b.wg * b.wg;
    const uint32_t i = (uint32_t)(g / wgsq), j = (uint32_t)(g % wgsq);
    rt_ray r;
    if (!camera_ray(b, i, j, r)) return;
    const uint32_t tilesX = b.width / b.wg;
    const uint32_t tileRow = i / tilesX;
    const uint64_t fullRows = b.height / b.wg;  // tile rows with wg pixel rows each
    uint64_t k;
    if (tileRow < fullRows) {
        k = g;
    } else {
        // partial last tile row: rows of the tile that exist
        const uint32_t rowsLeft = b.height - (uint32_t)fullRows * b.wg;
        k = fullRows * tilesX * wgsq + (uint64_t)(i - tileRow * tilesX) * rowsLeft * b.wg + j;
    }
    out[k] = r;
}

__global__ void __launch_bounds__(kShadeBlock) k_surface(const float* __restrict__ tris,
                                                         const uint8_t* __restrict__ leafLc, uint32_t Lm,
                                                         uint32_t LS, uint32_t nLeaves,
                                                         const rt_ray* __restrict__ rays,
                                                         const rt_hit* __restrict__ hits, uint32_t n,
                                                         float* __restrict__ nrmOut, float* __restrict__ posOut) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    if (i >= n) return;
    const rt_hit h = hits[i];
    float nrm[3], pos[3];
    const float qnan = __builtin_nanf("");
    const uint32_t ord = h.leaf / (9u * Lm);
    const uint32_t lc = ord < nLeaves ? leafLc[ord] : 0u;
    if (h.slot < 0 || lc == 0 || (uint32_t)h.slot >= lc || h.leaf % (9u * Lm) != 0) {
        nrm[0] = nrm[1] = nrm[2] = pos[0] = pos[1] = pos[2] = qnan;
    } else {
        // leaf block: p0, then e1 = p1 - p0 and e2 = p2 - p0 of the slot
        // (slots < lc are the leaf's triangles)
        const float* T = tris + (size_t)ord * LS;
        const uint32_t s = (uint32_t)h.slot;
        const float p0[3] = {T[leaf_offset(s, 0)], T[leaf_offset(s, 1)], T[leaf_offset(s, 2)]};
        const float e1[3] = {T[leaf_offset(s, 3)], T[leaf_offset(s, 4)], T[leaf_offset(s, 5)]};
        const float e2[3] = {T[leaf_offset(s, 6)], T[leaf_offset(s, 7)], T[leaf_offset(s, 8)]};
        surface(p0, e1, e2, rays[i], nrm, pos);
    }
    for (int a = 0; a < 3; ++a) {
        nrmOut[(size_t)i * 3 + a] = nrm[a];
        posOut[(size_t)i * 3 + a] = pos[a];
    }
}

// A ray that no box or triangle test accepts (tMax = -1 < every t0 >= 0), for
// the secondary ray of a pixel without a hit: one root visit, no hit, not
// occluded.  (A NaN origin, which the reference's stream any-hit skips,
// would pass every box test of the per-ray traversal: max/min drop NaN.)
__device__ rt_ray dead_ray() {
    const float o[3] = {0.0f, 0.0f, 0.0f}, t[3] = {0.0f, 0.0f, 1.0f};
    return make_ray(o, t, -1.0f);
}

// Dense shadow rays: ray i belongs to pixel i; a pixel without a hit gets
// dead_ray().
__global__ void __launch_bounds__(kShadeBlock) k_shadow(const float* __restrict__ nrm,
                                                        const float* __restrict__ pos,
                                                        const rt_hit* __restrict__ hits, uint32_t n, float lx,
                                                        float ly, float lz, float tmax, rt_ray* __restrict__ out) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    if (i >= n) return;
    const float light[3] = {lx, ly, lz};
    rt_ray r;
    if (hits[i].slot >= 0) {
        r = shadow_ray(nrm + (size_t)i * 3, pos + (size_t)i * 3, light, tmax);
    } else {
        r = dead_ray();
    }
    out[i] = r;
}

__global__ void k_zero_word(uint32_t* p) { *p = 0; }

struct IsHit {
    __host__ __device__ bool operator()(const rt_hit& h) const { return h.slot >= 0; }
};

// Diffuse ray r starts at hit pixel hitIdx[r % nHits] (the host's order);
// with no hit at all every ray is dead_ray().
__global__ void __launch_bounds__(kShadeBlock) k_diffuse(const rt_ray* __restrict__ primary,
                                                         const float* __restrict__ nrm,
                                                         const float* __restrict__ pos,
                                                         const uint32_t* __restrict__ hitIdx,
                                                         const uint32_t* __restrict__ nHitsPtr, uint64_t seed,
                                                         uint32_t nOut, float tmax, rt_ray* __restrict__ out) {
    const uint32_t r = blockIdx.x * kShadeBlock + threadIdx.x;
    if (r >= nOut) return;
    const uint32_t nHits = *nHitsPtr;
    rt_ray ray;
    if (nHits == 0) {
        ray = dead_ray();
    } else {
        const uint32_t i = hitIdx[r % nHits];
        const float pd[3] = {primary[i].dx, primary[i].dy, primary[i].dz};
        ray = diffuse_ray(pd, nrm + (size_t)i * 3, pos + (size_t)i * 3, seed, r, tmax);
    }
    out[r] = ray;
}

// Dense ambient rays of one sample: ray i = rt_spawn_ambient's ray for a hit
// pixel, dead_ray() for a miss (never occluded).
__global__ void __launch_bounds__(kShadeBlock) k_ambient(const float* __restrict__ nrm, const float* __restrict__ pos,
                                                         const rt_hit* __restrict__ hits, uint32_t n, uint32_t sample,
                                                         float dist, rt_ray* __restrict__ out) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = hits[i].slot >= 0 ? ambient_ray(i, sample, nrm + (size_t)i * 3, pos + (size_t)i * 3, dist) : dead_ray();
}

__global__ void __launch_bounds__(kShadeBlock) k_occl_add(const uint8_t* __restrict__ occ, uint32_t* __restrict__ acc,
                                                          uint32_t n) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    if (i < n) acc[i] += occ[i];
}

// Gray value per pixel (cameraFast.h:404-439): (uint8)((1 - occluded /
// samples) * 255) for a hit, 0 for a miss or with no samples.
__global__ void __launch_bounds__(kShadeBlock) k_gray(const rt_hit* __restrict__ hits,
                                                      const uint32_t* __restrict__ acc, uint32_t samples, uint32_t n,
                                                      uint8_t* __restrict__ gray) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    if (i >= n) return;
    uint8_t g = 0;
    if (hits[i].slot >= 0 && samples != 0) {
        const float factor = 1.0f - ((float)acc[i] / (float)samples);
        g = (uint8_t)(factor * 255.0f);
    }
    gray[i] = g;
}

}  // namespace

hipError_t launch_surface(const float* tris, const uint8_t* leafLc, uint32_t Lm, uint32_t nLeaves,
                          const rt_ray* rays, const rt_hit* hits, uint32_t n, float* nrm, float* pos,
                          hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_surface, dim3(blocks_for(n)), dim3(kShadeBlock), 0, stream, tris, leafLc, Lm,
                       leaf_stride(Lm), nLeaves, rays, hits, n, nrm, pos);
    return hipGetLastError();
}

}  // namespace rt

using namespace rt;

extern "C" int rt_camera_rays_device(const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                                     rt_ray* out, void* hip_stream) {
    if (!cam || !out || width == 0 || height == 0 || wg == 0)
        return set_error(RT_ERR_ARG, "rt_camera_rays_device: bad argument");
    if (width % wg != 0) return set_error(RT_ERR_ARG, "rt_camera_rays_device: width must be a multiple of wg");
    if ((uint64_t)width * height >= 0x80000000ull)
        return set_error(RT_ERR_ARG, "rt_camera_rays_device: at most 2^31-1 pixels");
    const CameraBasis b = camera_basis(*cam, width, height, wg);
    const uint64_t lanes = (uint64_t)(width / wg) * ((height + wg - 1) / wg) * wg * wg;
    hipLaunchKernelGGL(k_camera, dim3(blocks_for(lanes)), dim3(kShadeBlock), 0,
                       static_cast<hipStream_t>(hip_stream), b, lanes, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(RT_ERR_HIP, std::string("camera kernel: ") + hipGetErrorString(e));
    return RT_OK;
}

extern "C" int rt_spawn_shadow_device(const float* normal, const float* position, const rt_hit* hits, uint32_t n,
                                      const float light[3], float tmax, rt_ray* out, void* hip_stream) {
    if (!light || (n && (!normal || !position || !hits || !out)))
        return set_error(RT_ERR_ARG, "rt_spawn_shadow_device: bad argument");
    if (n == 0) return RT_OK;
    hipLaunchKernelGGL(k_shadow, dim3(blocks_for(n)), dim3(kShadeBlock), 0, static_cast<hipStream_t>(hip_stream),
                       normal, position, hits, n, light[0], light[1], light[2], tmax, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(RT_ERR_HIP, std::string("shadow kernel: ") + hipGetErrorString(e));
    return RT_OK;
}

extern "C" int rt_spawn_diffuse_device(const rt_ray* primary, const float* normal, const float* position,
                                       const rt_hit* hits, uint32_t n, uint64_t seed, uint32_t n_out, float tmax,
                                       rt_ray* out, void* hip_stream) {
    if (n_out && (!primary || !normal || !position || !hits || !out))
        return set_error(RT_ERR_ARG, "rt_spawn_diffuse_device: bad argument");
    if (n_out == 0) return RT_OK;
    if (n >= 0x80000000u) return set_error(RT_ERR_ARG, "rt_spawn_diffuse_device: at most 2^31-1 primary rays");
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    // ordered list of hit pixels (stream compaction), then one thread per ray
    size_t tempBytes = 0;
    hipcub::CountingInputIterator<uint32_t> idx(0);
    hipcub::TransformInputIterator<bool, IsHit, const rt_hit*> flags(hits, IsHit());
    hipError_t e = hipcub::DeviceSelect::Flagged(nullptr, tempBytes, idx, flags, (uint32_t*)nullptr,
                                                 (uint32_t*)nullptr, (int)n, st);
    if (e != hipSuccess) return set_error(RT_ERR_HIP, std::string("select sizing: ") + hipGetErrorString(e));
    const size_t listBytes = ((size_t)n * 4 + 255) / 256 * 256;
    char* buf = nullptr;
    e = hipMallocAsync(reinterpret_cast<void**>(&buf), listBytes + 256 + tempBytes, st);
    if (e != hipSuccess) return set_error(RT_ERR_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    uint32_t* hitIdx = reinterpret_cast<uint32_t*>(buf);
    uint32_t* nHits = reinterpret_cast<uint32_t*>(buf + listBytes);
    void* temp = buf + listBytes + 256;
    if (n == 0) {
        hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, st, nHits);
        e = hipGetLastError();
    } else {
        e = hipcub::DeviceSelect::Flagged(temp, tempBytes, idx, flags, hitIdx, nHits, (int)n, st);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_diffuse, dim3(blocks_for(n_out)), dim3(kShadeBlock), 0, st, primary, normal, position,
                           hitIdx, nHits, seed, n_out, tmax, out);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(buf, st);
    if (e != hipSuccess) return set_error(RT_ERR_HIP, std::string("diffuse spawn: ") + hipGetErrorString(e));
    if (f != hipSuccess) return set_error(RT_ERR_HIP, std::string("hipFreeAsync: ") + hipGetErrorString(f));
    return RT_OK;
}

extern "C" int rt_render_ambient(rt_scene_t s, const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                                 uint32_t flags, uint32_t samples, float dist, uint8_t* gray_out, float* ms_out,
                                 void* hip_stream) {
    if (!s || !cam || !gray_out || width == 0 || height == 0 || wg == 0)
        return set_error(RT_ERR_ARG, "rt_render_ambient: bad argument");
    if (flags & ~(RT_SAVE_DIST | RT_STREAM))
        return set_error(RT_ERR_ARG, "rt_render_ambient: flags are RT_CLOSEST or RT_SAVE_DIST, plus RT_STREAM");
    if (width % wg != 0) return set_error(RT_ERR_ARG, "rt_render_ambient: width must be a multiple of wg");
    const uint64_t n64 = (uint64_t)width * height;
    if (n64 >= 0x80000000ull) return set_error(RT_ERR_ARG, "rt_render_ambient: at most 2^31-1 pixels");
    const uint32_t n = (uint32_t)n64;
    // every step on one stream: the caller's, or the one rt_trace picks for
    // the scene when none is given (not the null stream, which the scene's
    // non-blocking stream does not wait for)
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    int prevDev = 0;
    (void)hipGetDevice(&prevDev);
    struct Restore {
        int prev;
        ~Restore() { (void)hipSetDevice(prev); }
    } restore{prevDev};
    if (!st) {
        const int r = scene_stream(s, &st);
        if (r != RT_OK) return r;
    }
    // rays | hits | normals | positions | ambient rays | occluded | counts | gray
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t oRays = 0, oHits = oRays + up(n * sizeof(rt_ray)), oNrm = oHits + up(n * sizeof(rt_hit));
    const size_t oPos = oNrm + up((size_t)n * 12), oAmb = oPos + up((size_t)n * 12);
    const size_t oOcc = oAmb + up(n * sizeof(rt_ray)), oAcc = oOcc + up(n), oGray = oAcc + up((size_t)n * 4);
    const size_t total = oGray + up(n);
    char* buf = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&buf), total, st);
    if (e != hipSuccess) return set_error(RT_ERR_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    rt_ray* rays = reinterpret_cast<rt_ray*>(buf + oRays);
    rt_hit* hits = reinterpret_cast<rt_hit*>(buf + oHits);
    float* nrm = reinterpret_cast<float*>(buf + oNrm);
    float* pos = reinterpret_cast<float*>(buf + oPos);
    rt_ray* amb = reinterpret_cast<rt_ray*>(buf + oAmb);
    uint8_t* occ = reinterpret_cast<uint8_t*>(buf + oOcc);
    uint32_t* acc = reinterpret_cast<uint32_t*>(buf + oAcc);
    uint8_t* gray = reinterpret_cast<uint8_t*>(buf + oGray);
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int rc = RT_OK;
    for (auto& x : ev)
        if (hipEventCreate(&x) != hipSuccess) rc = set_error(RT_ERR_HIP, "rt_render_ambient: hipEventCreate");
    auto step = [&](hipError_t err, const char* what) {
        if (rc == RT_OK && err != hipSuccess) rc = set_error(RT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    };
    if (rc == RT_OK) step(hipEventRecord(ev[0], st), "hipEventRecord");
    if (rc == RT_OK) rc = rt_camera_rays_device(cam, width, height, wg, rays, st);
    const uint32_t dev = RT_DEVICE_PTRS | RT_NO_SYNC;
    // camera rays are coherent: the exact-box records issue fewer instructions
    // for them (the quantized records, which move fewer bytes, serve the
    // ambient rays below)
    if (rc == RT_OK) rc = rt_trace(s, rays, n, flags | dev | RT_EXACT_BOXES, hits, nullptr, st);
    if (rc == RT_OK) step(hipEventRecord(ev[1], st), "hipEventRecord");
    if (rc == RT_OK && samples) {
        rc = rt_hit_surface_device(s, rays, hits, n, nrm, pos, st);
        if (rc == RT_OK) step(launch_zero(acc, n, st), "zero kernel");
        for (uint32_t a = 0; rc == RT_OK && a < samples; ++a) {
            hipLaunchKernelGGL(k_ambient, dim3(blocks_for(n)), dim3(kShadeBlock), 0, st, nrm, pos, hits, n, a, dist, amb);
            step(hipGetLastError(), "ambient kernel");
            if (rc == RT_OK) rc = rt_trace(s, amb, n, RT_ANY | (flags & RT_STREAM) | dev, occ, nullptr, st);
            if (rc == RT_OK) {
                hipLaunchKernelGGL(k_occl_add, dim3(blocks_for(n)), dim3(kShadeBlock), 0, st, occ, acc, n);
                step(hipGetLastError(), "occlusion kernel");
            }
        }
    }
    if (rc == RT_OK) {
        hipLaunchKernelGGL(k_gray, dim3(blocks_for(n)), dim3(kShadeBlock), 0, st, hits, acc, samples, n, gray);
        step(hipGetLastError(), "gray kernel");
    }
    if (rc == RT_OK) step(hipEventRecord(ev[2], st), "hipEventRecord");
    if (rc == RT_OK) step(hipMemcpyAsync(gray_out, gray, n, hipMemcpyDeviceToHost, st), "hipMemcpy(gray)");
    const hipError_t f = hipFreeAsync(buf, st);
    step(hipStreamSynchronize(st), "rt_render_ambient");
    step(f, "hipFreeAsync");
    if (rc == RT_OK && ms_out) {
        step(hipEventElapsedTime(&ms_out[0], ev[0], ev[1]), "hipEventElapsedTime");
        step(hipEventElapsedTime(&ms_out[1], ev[0], ev[2]), "hipEventElapsedTime");
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    return rc;
}
