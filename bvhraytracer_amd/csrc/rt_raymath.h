// Ray construction and hit shading arithmetic shared by the host code (g++)
// and the device kernels (hipcc).  Both sides compile this one source with
// -ffp-contract=off, and division and sqrt are correctly rounded on both, so
// rays, normals and positions made on the GPU are bit-identical to the host's.
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/rt.h"

#ifdef __HIPCC__
#define RT_HD __host__ __device__ inline
#else
#define RT_HD inline
#endif

namespace rt {

// FastRay::FastRay (fastRay.cpp:3-20): dir = target - pos (not normalised),
// inv = 1/dir with FLT_MAX where a component is 0.
RT_HD void set_direction(rt_ray& r, float dx, float dy, float dz) {
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    r.ix = (dx == 0.0f) ? FLT_MAX : 1.0f / dx;
    r.iy = (dy == 0.0f) ? FLT_MAX : 1.0f / dy;
    r.iz = (dz == 0.0f) ? FLT_MAX : 1.0f / dz;
}

RT_HD rt_ray make_ray(const float o[3], const float target[3], float tmax) {
    rt_ray r;
    r.tmax = tmax;
    r.ox = o[0];
    r.oy = o[1];
    r.oz = o[2];
    set_direction(r, target[0] - o[0], target[1] - o[1], target[2] - o[2]);
    return r;
}

// glm::cross / glm::dot / glm::normalize (v * (1 / sqrt(dot(v, v))))
RT_HD void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
RT_HD float dot3(const float a[3], const float b[3]) {
    const float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
    return x + y + z;
}
RT_HD void normalize3(float v[3]) {
    const float s = 1.0f / sqrtf(dot3(v, v));
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
}

// Camera basis of inverse(lookAt(pos, target, up)) (camera.h:52-58) in
// closed form: columns s, u, -f, pos.
struct CameraBasis {
    float s[3], u[3], nf[3], pos[3];
    float focal, invW;
    uint32_t width, height, wg;
};

RT_HD CameraBasis camera_basis(const rt_camera& cam, uint32_t width, uint32_t height, uint32_t wg) {
    CameraBasis b;
    float f[3] = {cam.target[0] - cam.pos[0], cam.target[1] - cam.pos[1], cam.target[2] - cam.pos[2]};
    normalize3(f);
    cross3(f, cam.up, b.s);
    normalize3(b.s);
    cross3(b.s, f, b.u);
    for (int a = 0; a < 3; ++a) {
        b.nf[a] = -f[a];
        b.pos[a] = cam.pos[a];
    }
    b.focal = cam.focal;
    b.invW = 1.0f / (float)width;
    b.width = width;
    b.height = height;
    b.wg = wg;
    return b;
}

// Ray of lane j of tile i (CameraFast tile order, cameraFast.h:318-342 and
// RenderInfo, :324-327), or false when the pixel lies below the image (the
// partial last tile row).
RT_HD bool camera_ray(const CameraBasis& b, uint32_t i, uint32_t j, rt_ray& out) {
    const uint32_t wg = b.wg, width = b.width, height = b.height;
    const uint32_t row = (i * wg) / width * wg + j / wg;
    if (row >= height) return false;
    const int w = (int)(((float)((i * wg) % width) - width / 2.f) + (float)(j % wg));
    const int h = (int)(-((float)(((i * wg) / width) * wg) - height / 2.f) - (float)(j / wg));
    // Camera::getRayTargetPosition (camera.h:91-98):
    // centerOffset = (w, h, 0, 0) * (1/width) + (0, 0, -focal, 0)
    const float cx = (float)w * b.invW, cy = (float)h * b.invW, cz = 0.0f * b.invW + -b.focal;
    float tgt[3];
    for (int a = 0; a < 3; ++a) {
        // glm mat4*vec4: (m0*v0 + m1*v1) + (m2*v2 + m3*v3), v3 = 0
        const float add0 = b.s[a] * cx + b.u[a] * cy;
        const float add1 = b.nf[a] * cz + b.pos[a] * 0.0f;
        tgt[a] = b.pos[a] + (add0 + add1);
    }
    out = make_ray(b.pos, tgt, 222222.0f);
    return true;
}

// getSurfaceNormalPosition (fastNodeManager.h:1699-1726) from p0 and the
// edges e1 = p1 - p0, e2 = p2 - p0: n = normalize(cross(e1, e2)) (computeNormal,
// glmUtil.h:36-39); the position re-solves the plane, t' = dot(p0 - o, n) /
// dot(n, dir), P = o + dir * t'.
RT_HD void surface(const float p0[3], const float e1[3], const float e2[3], const rt_ray& r, float nrm[3],
                   float pos[3]) {
    cross3(e1, e2, nrm);
    normalize3(nrm);
    const float dir[3] = {r.dx, r.dy, r.dz}, o[3] = {r.ox, r.oy, r.oz};
    const float denom = dot3(nrm, dir);
    const float p0l0[3] = {p0[0] - o[0], p0[1] - o[1], p0[2] - o[2]};
    const float t = dot3(p0l0, nrm) / denom;
    for (int a = 0; a < 3; ++a) pos[a] = o[a] + dir[a] * t;
}

// Shadow ray toward a point light: FastRay(P + n*0.001, light, tmax)
// (cameraFast.h:373,521; unnormalised direction, so t in [0,1] reaches the
// light).
RT_HD rt_ray shadow_ray(const float nrm[3], const float pos[3], const float light[3], float tmax) {
    float o[3];
    for (int a = 0; a < 3; ++a) o[a] = pos[a] + nrm[a] * 0.001f;
    return make_ray(o, light, tmax);
}

// PCG32 (O'Neill, PCG-XSH-RR 64/32) step and seeding.
struct Pcg {
    uint64_t state, inc;
};
RT_HD uint32_t pcg_next(Pcg& g) {
    const uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}
RT_HD Pcg pcg_seed(uint64_t seed, uint64_t stream) {
    Pcg g;
    g.state = 0;
    g.inc = (stream << 1u) | 1u;
    pcg_next(g);
    g.state += seed;
    pcg_next(g);
    return g;
}
// [0,1) with 24 bits: (u >> 8) * 2^-24
RT_HD float pcg_uniform(Pcg& g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }

// sin and cos of 2*pi*u for u in [0, 1), the same bits on host and device.
// The angle is reduced in turns, which is exact: quadrant q = floor(4u),
// remainder g = 4u - q in [0, 1), folded to [0, 1/2] (swapping sin and cos),
// x = g * pi/2 in [0, pi/4].  sin and cos of x come from their Taylor series
// through x^9 and x^10 (truncation error below 2e-9 on [0, pi/4]), evaluated
// by Horner's rule; the quadrant then rotates the pair.
RT_HD void sincos_turn(float u, float* s, float* c) {
    const float u4 = u * 4.0f;
    int q = (int)u4;
    float g = u4 - (float)q;
    const bool fold = g > 0.5f;
    if (fold) g = 1.0f - g;
    const float x = g * 1.57079632679489662f;
    const float x2 = x * x;
    const float sx =
        x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
    const float cx =
        1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f +
                                   x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f + x2 * (-1.0f / 3628800.0f)))));
    float sn = fold ? cx : sx, cs = fold ? sx : cx;
    q &= 3;
    if (q == 0) {
        *s = sn;
        *c = cs;
    } else if (q == 1) {
        *s = cs;
        *c = -sn;
    } else if (q == 2) {
        *s = -sn;
        *c = -cs;
    } else {
        *s = -cs;
        *c = sn;
    }
}

// Cosine-distributed bounce ray number r (PCG stream r of `seed`) from a
// primary hit: the normal is face-forwarded against the primary direction,
// the direction is Malley's method (a uniform disk point lifted to the
// hemisphere) in the orthonormal basis of Duff et al. 2017, and the origin
// is P + n*0.001 as for the reference's secondary rays.
RT_HD rt_ray diffuse_ray(const float primaryDir[3], const float normal[3], const float pos[3], uint64_t seed,
                         uint64_t r, float tmax) {
    float nn[3] = {normal[0], normal[1], normal[2]};
    if (dot3(nn, primaryDir) > 0.0f) {
        nn[0] = -nn[0];
        nn[1] = -nn[1];
        nn[2] = -nn[2];
    }
    Pcg g = pcg_seed(seed, r);
    const float u1 = pcg_uniform(g), u2 = pcg_uniform(g);
    float sp, cp;
    sincos_turn(u2, &sp, &cp);
    const float rad = sqrtf(u1);
    const float lx = rad * cp, ly = rad * sp, lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
    const float sgn = copysignf(1.0f, nn[2]);
    const float a = -1.0f / (sgn + nn[2]);
    const float bb = nn[0] * nn[1] * a;
    const float t[3] = {1.0f + sgn * nn[0] * nn[0] * a, sgn * bb, -sgn * nn[0]};
    const float bt[3] = {bb, sgn + nn[1] * nn[1] * a, -nn[1]};
    rt_ray ray;
    ray.tmax = tmax;
    ray.ox = pos[0] + nn[0] * 0.001f;
    ray.oy = pos[1] + nn[1] * 0.001f;
    ray.oz = pos[2] + nn[2] * 0.001f;
    set_direction(ray, t[0] * lx + bt[0] * ly + nn[0] * lz, t[1] * lx + bt[1] * ly + nn[1] * lz,
                  t[2] * lx + bt[2] * ly + nn[2] * lz);
    return ray;
}

// Ambient-occlusion direction number `sample` of pixel `index`, the role of
// Camera::getAmbientDirection + sampleHemisphere (camera.cpp:21-52): a
// direction uniform on the unit sphere, negated when its dot product with
// the surface normal is <= 0.  The reference hashes (index, sample) with
// std::hash, whose value is implementation defined; here (u, v) come from a
// fixed 64-bit mix of index and sample (SplitMix64's finaliser constants),
// 24 bits each.  The sphere point is z = 1 - 2u, r = sqrt(1 - z^2),
// (x, y) = r (cos 2 pi v, sin 2 pi v), the same distribution as the
// reference's phi = acos(1 - 2u), theta = 2 pi v form.
RT_HD void ambient_dir(uint64_t index, uint32_t sample, const float normal[3], float d[3]) {
    uint64_t h = index * 0x9E3779B97F4A7C15ull + (uint64_t)sample * 0xBF58476D1CE4E5B9ull + 1ull;
    h ^= h >> 31;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 29;
    const float u = (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
    const float v = (float)(uint32_t)((h >> 16) & 0xFFFFFFull) * (1.0f / 16777216.0f);
    const float z = 1.0f - 2.0f * u;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sv, cv;
    sincos_turn(v, &sv, &cv);
    d[0] = r * cv;
    d[1] = r * sv;
    d[2] = z;
    if (dot3(d, normal) <= 0.0f) {
        d[0] = -d[0];
        d[1] = -d[1];
        d[2] = -d[2];
    }
}

// Ambient ray of the performance renderer (cameraFast.h:395-418): origin
// P + n*0.001, FastRay::updateDirection with the ambient direction itself
// (fastRay.cpp:22-38), tMax = ambientDistance.
RT_HD rt_ray ambient_ray(uint64_t index, uint32_t sample, const float nrm[3], const float pos[3], float dist) {
    float d[3];
    ambient_dir(index, sample, nrm, d);
    rt_ray r;
    r.tmax = dist;
    r.ox = pos[0] + nrm[0] * 0.001f;
    r.oy = pos[1] + nrm[1] * 0.001f;
    r.oz = pos[2] + nrm[2] * 0.001f;
    set_direction(r, d[0], d[1], d[2]);
    return r;
}

}  // namespace rt
