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

namespace rt {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int build_bvh(const float* tris, uint32_t n, const rt_build_params& p, HostBvh& out);

namespace {

// FastRay::FastRay (fastRay.cpp:3-20): dir = target - pos (not normalised),
// inv = 1/dir with FLT_MAX where a component is 0.
rt_ray make_ray(const float o[3], const float target[3], float tmax) {
    rt_ray r;
    r.tmax = tmax;
    r.ox = o[0];
    r.oy = o[1];
    r.oz = o[2];
    r.dx = target[0] - o[0];
    r.dy = target[1] - o[1];
    r.dz = target[2] - o[2];
    r.ix = (r.dx == 0.0f) ? FLT_MAX : 1.0f / r.dx;
    r.iy = (r.dy == 0.0f) ? FLT_MAX : 1.0f / r.dy;
    r.iz = (r.dz == 0.0f) ? FLT_MAX : 1.0f / r.dz;
    return r;
}

// FastRay::updateDirection (fastRay.cpp:22-38)
void set_direction(rt_ray& r, float dx, float dy, float dz) {
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    r.ix = (dx == 0.0f) ? FLT_MAX : 1.0f / dx;
    r.iy = (dy == 0.0f) ? FLT_MAX : 1.0f / dy;
    r.iz = (dz == 0.0f) ? FLT_MAX : 1.0f / dz;
}

inline void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
inline float dot3(const float a[3], const float b[3]) {
    float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
    return x + y + z;
}
// glm::normalize: v * (1 / sqrt(dot(v, v)))
inline void normalize3(float v[3]) {
    float s = 1.0f / std::sqrt(dot3(v, v));
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
}

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
    float f[3] = {cam->target[0] - cam->pos[0], cam->target[1] - cam->pos[1], cam->target[2] - cam->pos[2]};
    normalize3(f);
    float s[3];
    cross3(f, cam->up, s);
    normalize3(s);
    float u[3];
    cross3(s, f, u);
    const float nf[3] = {-f[0], -f[1], -f[2]};
    const float invW = 1.0f / (float)width;
    const uint32_t tilesX = width / wg;
    const uint32_t tileRows = (height + wg - 1) / wg;
    const uint32_t wgsq = wg * wg;
    size_t k = 0;
    for (uint32_t i = 0; i < tilesX * tileRows; ++i) {
        for (uint32_t j = 0; j < wgsq; ++j) {
            const uint32_t row = (i * wg) / width * wg + j / wg;
            if (row >= height) continue;  // partial last tile row
            // RenderInfo(w, h) of cameraFast.h:324-327 (float -> int)
            const int w = (int)(((float)((i * wg) % width) - width / 2.f) + (float)(j % wg));
            const int h = (int)(-((float)(((i * wg) / width) * wg) - height / 2.f) - (float)(j / wg));
            // centerOffset = (w, h, 0, 0) * (1/width) + (0, 0, -focal, 0)
            const float cx = (float)w * invW, cy = (float)h * invW, cz = 0.0f * invW + -cam->focal;
            float tgt[3];
            for (int a = 0; a < 3; ++a) {
                // glm mat4*vec4: (m0*v0 + m1*v1) + (m2*v2 + m3*v3), v3 = 0
                const float add0 = s[a] * cx + u[a] * cy;
                const float add1 = nf[a] * cz + cam->pos[a] * 0.0f;
                tgt[a] = cam->pos[a] + (add0 + add1);
            }
            out[k++] = make_ray(cam->pos, tgt, 222222.0f);
        }
    }
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
        float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float nrm[3];
        cross3(e1, e2, nrm);
        normalize3(nrm);
        const rt_ray& r = rays[i];
        const float dir[3] = {r.dx, r.dy, r.dz}, o[3] = {r.ox, r.oy, r.oz};
        const float denom = dot3(nrm, dir);
        const float p0l0[3] = {p0[0] - o[0], p0[1] - o[1], p0[2] - o[2]};
        const float t = dot3(p0l0, nrm) / denom;
        for (int a = 0; a < 3; ++a) {
            nn[a] = nrm[a];
            pp[a] = o[a] + dir[a] * t;
        }
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
        const float* nn = normal + (size_t)i * 3;
        const float* pp = position + (size_t)i * 3;
        float o[3];
        for (int a = 0; a < 3; ++a) o[a] = pp[a] + nn[a] * 0.001f;
        out[k] = make_ray(o, light, tmax);
        if (src_index) src_index[k] = i;
        ++k;
    }
    *n_out = k;
    return RT_OK;
}

// Cosine-distributed bounce rays (Malley's method, orthonormal basis of Duff
// et al. 2017) around the hit normal face-forwarded against the primary ray.
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
        const float* pp = position + (size_t)i * 3;
        float nn[3] = {normal[i * 3 + 0], normal[i * 3 + 1], normal[i * 3 + 2]};
        const float pd[3] = {primary[i].dx, primary[i].dy, primary[i].dz};
        if (dot3(nn, pd) > 0.0f) {
            nn[0] = -nn[0];
            nn[1] = -nn[1];
            nn[2] = -nn[2];
        }
        Pcg32 rng(seed, r);
        const float u1 = rng.uniform(), u2 = rng.uniform();
        const float rad = std::sqrt(u1), phi = 6.2831853071795865f * u2;
        const float lx = rad * std::cos(phi), ly = rad * std::sin(phi), lz = std::sqrt(std::max(0.0f, 1.0f - u1));
        const float sgn = std::copysign(1.0f, nn[2]);
        const float a = -1.0f / (sgn + nn[2]);
        const float bb = nn[0] * nn[1] * a;
        const float t[3] = {1.0f + sgn * nn[0] * nn[0] * a, sgn * bb, -sgn * nn[0]};
        const float bt[3] = {bb, sgn + nn[1] * nn[1] * a, -nn[1]};
        rt_ray ray;
        ray.tmax = tmax;
        ray.ox = pp[0] + nn[0] * 0.001f;
        ray.oy = pp[1] + nn[1] * 0.001f;
        ray.oz = pp[2] + nn[2] * 0.001f;
        set_direction(ray, t[0] * lx + bt[0] * ly + nn[0] * lz, t[1] * lx + bt[1] * ly + nn[1] * lz,
                      t[2] * lx + bt[2] * ly + nn[2] * lz);
        out[r] = ray;
    }
    return RT_OK;
}
