// Host-side C ABI: builder, camera, hit reconstruction, secondary rays,
// SAH-factor table and error reporting.
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>

#include "host_common.h"
#include "rt_raymath.h"

namespace rt {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int build_bvh(const float* tris, uint32_t n, const rt_build_params& p, HostBvh& out);

namespace {

// sahFactorConfigSse.txt:2-17 — (branching, leaf size) -> SAH node factor.
struct SahRow {
    int n, l;
    float f;
};
const SahRow kSahSse[16] = {
    {4, 4, 0.9790158394920848f},   {4, 8, 0.6019200777645688f},   {4, 12, 0.46148535855892625f},
    {4, 16, 0.38099865799955585f}, {8, 4, 1.3129201663140904f},   {8, 8, 0.8089449478051558f},
    {8, 12, 0.6356877674077516f},  {8, 16, 0.5431092069456159f},  {12, 4, 1.5611955563121058f},
    {12, 8, 0.9667477491937599f},  {12, 12, 0.7477536713402126f}, {12, 16, 0.638505587529224f},
    {16, 4, 1.8033700927282266f},  {16, 8, 1.1369639789217003f},  {16, 12, 0.8735301252614894f},
    {16, 16, 0.7655054621988963f}};

}  // namespace
}  // namespace rt

using namespace rt;

struct rt_host_bvh {
    HostBvh b;
};

extern "C" const char* rt_last_error(void) { return g_last_error.c_str(); }

extern "C" int rt_build(const float* tris, uint32_t n_tris, const rt_build_params* p,
                        rt_host_bvh_t* out) {
    if (!p || !out) return set_error(RT_ERR_ARG, "rt_build: null argument");
    rt_host_bvh* h = new (std::nothrow) rt_host_bvh;
    if (!h) return set_error(RT_ERR_NOMEM, "rt_build: out of host memory");
    int rc;
    try {
        rc = build_bvh(tris, n_tris, *p, h->b);
    } catch (const std::bad_alloc&) {
        rc = set_error(RT_ERR_NOMEM, "rt_build: out of host memory");
    }
    if (rc != RT_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return RT_OK;
}

extern "C" int rt_host_bvh_desc(rt_host_bvh_t h, rt_bvh_desc* d) {
    if (!h || !d) return set_error(RT_ERR_ARG, "rt_host_bvh_desc: null argument");
    const HostBvh& b = h->b;
    d->branching = b.N;
    d->leaf_size = b.L;
    d->gang = b.G;
    d->node_memory = b.M;
    d->leaf_memory = b.Lm;
    d->node_count = b.nodeCount();
    d->bounds = b.bounds.data();
    d->base = b.base.data();
    d->order = b.order.data();
    d->child_type = b.childType.data();
    d->tri_float_count = b.tri.size();
    d->tri_points = b.tri.data();
    d->tri_ids = b.triIds.data();
    d->average_depth = b.averageDepth;
    d->triangle_count = b.triangleCount;
    d->max_depth = b.maxDepth;
    return RT_OK;
}

extern "C" int rt_host_bvh_free(rt_host_bvh_t h) {
    delete h;
    return RT_OK;
}

// RayTracer::loadSahFactor (rayTracer.cpp:880-925): gang 8 -> 1 (the AVX
// table is not set up); lookup by (roundup(N,G), roundup(L,G)); not found -> 1.
// table_path (optional) names a file in the sahFactorConfigSse.txt format.
extern "C" float rt_sah_factor(uint32_t leaf_size, uint32_t branching, uint32_t gang,
                               const char* table_path) {
    if (gang == 8) return 1.0f;
    const int g = (int)(gang ? gang : 4);
    const int tl = (int)std::ceil(leaf_size / (float)g) * g;
    const int tn = (int)std::ceil(branching / (float)g) * g;
    if (table_path && *table_path) {
        std::ifstream f(table_path);
        std::string line;
        if (f.is_open() && std::getline(f, line)) {
            while (std::getline(f, line)) {
                if (line.empty()) continue;
                int cn = 0, cl = 0;
                float fac = NAN;
                if (std::sscanf(line.c_str(), "%d, %d, %f", &cn, &cl, &fac) == 3 && cn == tn && cl == tl)
                    return fac;
            }
        }
        return 1.0f;
    }
    for (const auto& r : kSahSse)
        if (r.n == tn && r.l == tl) return r.f;
    return 1.0f;
}

// Camera::getRayTargetPosition (camera.h:91-98) with the camera-to-world
// transform inverse(lookAt(pos, target, up)) (camera.h:52-58) written in
// closed form: columns s, u, -f, pos.  The reference obtains it with
// glm::inverse; the closed form can differ from it in the last bit.
extern "C" int rt_camera_rays(const rt_camera* cam, uint32_t width, uint32_t height, uint32_t wg,
                              rt_ray* out) {
    if (!cam || !out || width == 0 || height == 0 || wg == 0)
        return set_error(RT_ERR_ARG, "rt_camera_rays: bad argument");
    if (width % wg != 0) return set_error(RT_ERR_ARG, "rt_camera_rays: width must be a multiple of wg");
    const CameraBasis b = camera_basis(*cam, width, height, wg);
    const uint32_t tiles = width / wg * ((height + wg - 1) / wg);
    size_t k = 0;
    for (uint32_t i = 0; i < tiles; ++i)
        for (uint32_t j = 0; j < wg * wg; ++j)
            if (camera_ray(b, i, j, out[k])) ++k;
    return RT_OK;
}

// FastNodeManager::getSurfaceNormalPosition / getTriPoints / getLeafLoopCount
// (fastNodeManager.h:1699-1746), computeNormal (glmUtil.h:36-39).
extern "C" int rt_hit_surface(rt_host_bvh_t hb, const rt_ray* rays, const rt_hit* hits, uint32_t n,
                              float* normal_out, float* position_out) {
    if (!hb || (n && (!rays || !hits || !normal_out || !position_out)))
        return set_error(RT_ERR_ARG, "rt_hit_surface: bad argument");
    const HostBvh& b = hb->b;
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    for (uint32_t i = 0; i < n; ++i) {
        float* nn = normal_out + (size_t)i * 3;
        float* pp = position_out + (size_t)i * 3;
        if (hits[i].slot < 0) {
            nn[0] = nn[1] = nn[2] = pp[0] = pp[1] = pp[2] = qnan;
            continue;
        }
        const uint32_t leaf = hits[i].leaf;
        if ((size_t)leaf + 9 * b.Lm > b.tri.size() || hits[i].slot >= (int)b.Lm)
            return set_error(RT_ERR_ARG, "rt_hit_surface: hit outside the triangle array");
        int lc = (int)b.Lm;
        if (b.Lm != b.G) {
            const size_t end = leaf + (size_t)b.Lm * 9;
            if (std::isnan(b.tri[end - 1])) lc = (int)b.tri[end - 2];
        }
        const size_t t0 = (size_t)leaf + hits[i].slot;
        float p0[3], p1[3], p2[3];
        for (int a = 0; a < 3; ++a) {
            p0[a] = b.tri[t0 + (size_t)lc * a];
            p1[a] = b.tri[t0 + (size_t)lc * (3 + a)];
            p2[a] = b.tri[t0 + (size_t)lc * (6 + a)];
        }
        const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        surface(p0, e1, e2, rays[i], nn, pp);
    }
    return RT_OK;
}

// Secondary origin P + n*0.001 (cameraFast.h:373,521) and FastRay(origin,
// light, tmax): unnormalised direction, so t in [0,1] reaches the light.
extern "C" int rt_spawn_shadow(const float* normal, const float* position, const rt_hit* hits,
                               uint32_t n, const float light[3], float tmax, rt_ray* out,
                               uint32_t* src_index, uint32_t* n_out) {
    if (!light || !n_out || (n && (!normal || !position || !hits || !out)))
        return set_error(RT_ERR_ARG, "rt_spawn_shadow: bad argument");
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (hits[i].slot < 0) continue;
        out[k] = shadow_ray(normal + (size_t)i * 3, position + (size_t)i * 3, light, tmax);
        if (src_index) src_index[k] = i;
        ++k;
    }
    *n_out = k;
    return RT_OK;
}

// Ambient rays (ambient_ray) of one sample for every hit pixel, compacted.
extern "C" int rt_spawn_ambient(const float* normal, const float* position, const rt_hit* hits, uint32_t n,
                                uint32_t sample, float dist, rt_ray* out, uint32_t* src_index, uint32_t* n_out) {
    if (!n_out || (n && (!normal || !position || !hits || !out)))
        return set_error(RT_ERR_ARG, "rt_spawn_ambient: bad argument");
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (hits[i].slot < 0) continue;
        out[k] = ambient_ray(i, sample, normal + (size_t)i * 3, position + (size_t)i * 3, dist);
        if (src_index) src_index[k] = i;
        ++k;
    }
    *n_out = k;
    return RT_OK;
}

// Cosine-distributed bounce rays (diffuse_ray) cycling over the hit pixels:
// ray r starts at hit number r % hits.
extern "C" int rt_spawn_diffuse(const rt_ray* primary, const float* normal, const float* position,
                                const rt_hit* hits, uint32_t n, uint64_t seed, uint32_t n_out,
                                float tmax, rt_ray* out) {
    if (n_out && (!primary || !normal || !position || !hits || !out))
        return set_error(RT_ERR_ARG, "rt_spawn_diffuse: bad argument");
    std::vector<uint32_t> hitIdx;
    for (uint32_t i = 0; i < n; ++i)
        if (hits[i].slot >= 0) hitIdx.push_back(i);
    if (hitIdx.empty() && n_out) return set_error(RT_ERR_ARG, "rt_spawn_diffuse: no hits to spawn from");
    for (uint32_t r = 0; r < n_out; ++r) {
        const uint32_t i = hitIdx[r % hitIdx.size()];
        const float pd[3] = {primary[i].dx, primary[i].dy, primary[i].dz};
        out[r] = diffuse_ray(pd, normal + (size_t)i * 3, position + (size_t)i * 3, seed, r, tmax);
    }
    return RT_OK;
}
