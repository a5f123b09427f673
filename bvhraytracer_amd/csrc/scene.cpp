// Procedural scenes for the benchmark configurations.  The reference loads
// glTF models (src/modelLoader.h:23); the workloads here are synthetic and
// seeded, and keep the reference's zero-area filter (gameobject.h:120-122).
#include <algorithm>
#include <cstdlib>

#include "host_common.h"

namespace rt {

namespace {

struct D3 {
    double x, y, z;
};

inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 dadd(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 dmul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 dnorm(D3 a) {
    double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

// calcTriangleSurfaceArea (glmUtil.h:56-61) != 0, in FP32 like the reference.
bool nonzero_area(const float* t) {
    float ax = t[3] - t[0], ay = t[4] - t[1], az = t[5] - t[2];
    float bx = t[6] - t[0], by = t[7] - t[1], bz = t[8] - t[2];
    float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    float len = std::sqrt(cx * cx + cy * cy + cz * cz);
    return 0.5f * len != 0.0f;
}

void emit(std::vector<float>& out, D3 a, D3 b, D3 c) {
    float t[9] = {(float)a.x, (float)a.y, (float)a.z, (float)b.x, (float)b.y,
                  (float)b.z, (float)c.x, (float)c.y, (float)c.z};
    if (nonzero_area(t)) out.insert(out.end(), t, t + 9);
}

const double kPhi = 1.6180339887498949;
const D3 kIcoV[12] = {{-1, kPhi, 0}, {1, kPhi, 0}, {-1, -kPhi, 0}, {1, -kPhi, 0},
                      {0, -1, kPhi}, {0, 1, kPhi}, {0, -1, -kPhi}, {0, 1, -kPhi},
                      {kPhi, 0, -1}, {kPhi, 0, 1}, {-kPhi, 0, -1}, {-kPhi, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                          {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                          {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                          {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};

// Geodesic icosphere of frequency f: every icosahedron face is cut into f*f
// triangles and the grid points are projected onto the sphere (20*f^2 tris).
void icosphere(std::vector<float>& out, D3 c, double r, int f) {
    for (int fi = 0; fi < 20; ++fi) {
        D3 A = dnorm(kIcoV[kIcoF[fi][0]]), B = dnorm(kIcoV[kIcoF[fi][1]]),
           C = dnorm(kIcoV[kIcoF[fi][2]]);
        auto P = [&](int i, int j) {
            D3 p = dadd(A, dadd(dmul(dsub(B, A), (double)i / f), dmul(dsub(C, A), (double)j / f)));
            return dadd(c, dmul(dnorm(p), r));
        };
        for (int i = 0; i < f; ++i)
            for (int j = 0; j < f - i; ++j) {
                emit(out, P(i, j), P(i + 1, j), P(i, j + 1));
                if (i + j < f - 1) emit(out, P(i + 1, j), P(i + 1, j + 1), P(i, j + 1));
            }
    }
}

void quad(std::vector<float>& out, D3 a, D3 b, D3 c, D3 d) {
    emit(out, a, b, c);
    emit(out, a, c, d);
}

}  // namespace

// Spheres scene: `count` icospheres of frequency `freq`, radius U[0.5,2],
// centre U[-10,10]^3, plus a 2-triangle ground plane y=-12 over [-40,40]^2.
std::vector<float> scene_spheres(uint64_t seed, uint32_t count, uint32_t freq) {
    Pcg32 rng(seed, 1);
    std::vector<float> out;
    out.reserve((size_t)count * 20 * freq * freq * 9 + 18);
    for (uint32_t s = 0; s < count; ++s) {
        double r = 0.5 + 1.5 * rng.uniform_d();
        D3 c;
        c.x = -10.0 + 20.0 * rng.uniform_d();
        c.y = -10.0 + 20.0 * rng.uniform_d();
        c.z = -10.0 + 20.0 * rng.uniform_d();
        icosphere(out, c, r, (int)freq);
    }
    quad(out, {-40, -12, -40}, {-40, -12, 40}, {40, -12, 40}, {40, -12, -40});
    return out;
}

// Interior scene: 12-triangle room shell [-15,15]x[0,10]x[-15,15] facing
// inward, plus (count-12) triangles in `clusters` Gaussian clusters
// (sigma 0.3, centroids clamped to the room), uniformly random orientation,
// edge length e*(1 +- 0.25) with e ~ LogNormal(-3, 0.7).
std::vector<float> scene_interior(uint64_t seed, uint32_t count, uint32_t clusters) {
    Pcg32 rng(seed, 2);
    std::vector<float> out;
    out.reserve((size_t)count * 9);
    const double X0 = -15, X1 = 15, Y0 = 0, Y1 = 10, Z0 = -15, Z1 = 15;
    quad(out, {X0, Y0, Z0}, {X1, Y0, Z0}, {X1, Y0, Z1}, {X0, Y0, Z1});  // floor
    quad(out, {X0, Y1, Z0}, {X0, Y1, Z1}, {X1, Y1, Z1}, {X1, Y1, Z0});  // ceiling
    quad(out, {X0, Y0, Z0}, {X0, Y0, Z1}, {X0, Y1, Z1}, {X0, Y1, Z0});  // -x wall
    quad(out, {X1, Y0, Z0}, {X1, Y1, Z0}, {X1, Y1, Z1}, {X1, Y0, Z1});  // +x wall
    quad(out, {X0, Y0, Z0}, {X0, Y1, Z0}, {X1, Y1, Z0}, {X1, Y0, Z0});  // -z wall
    quad(out, {X0, Y0, Z1}, {X1, Y0, Z1}, {X1, Y1, Z1}, {X0, Y1, Z1});  // +z wall
    if (count <= 12) return out;
    uint32_t nc = clusters ? clusters : std::max<uint32_t>(1, (uint32_t)((uint64_t)count * 500 / 1000000));
    std::vector<D3> cen(nc);
    for (uint32_t i = 0; i < nc; ++i) {
        cen[i].x = X0 + (X1 - X0) * rng.uniform_d();
        cen[i].y = Y0 + (Y1 - Y0) * rng.uniform_d();
        cen[i].z = Z0 + (Z1 - Z0) * rng.uniform_d();
    }
    const uint32_t body = count - 12;
    for (uint32_t i = 0; i < body; ++i) {
        const D3& cc = cen[i % nc];
        D3 p;
        p.x = std::min(X1, std::max(X0, cc.x + 0.3 * rng.normal()));
        p.y = std::min(Y1, std::max(Y0, cc.y + 0.3 * rng.normal()));
        p.z = std::min(Z1, std::max(Z0, cc.z + 0.3 * rng.normal()));
        double e = std::exp(-3.0 + 0.7 * rng.normal());
        // uniform random rotation (Shoemake's unit quaternion)
        double u1 = rng.uniform_d(), u2 = rng.uniform_d(), u3 = rng.uniform_d();
        double qa = std::sqrt(1 - u1) * std::sin(6.283185307179586 * u2);
        double qb = std::sqrt(1 - u1) * std::cos(6.283185307179586 * u2);
        double qc = std::sqrt(u1) * std::sin(6.283185307179586 * u3);
        double qd = std::sqrt(u1) * std::cos(6.283185307179586 * u3);
        // rotation matrix of quaternion (w=qd, x=qa, y=qb, z=qc)
        double w = qd, x = qa, y = qb, z = qc;
        double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                          {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                          {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
        D3 v[3];
        for (int k = 0; k < 3; ++k) {
            double rho = e / std::sqrt(3.0) * (1.0 + 0.25 * (2.0 * rng.uniform_d() - 1.0));
            double ang = 6.283185307179586 * k / 3.0;
            double lx = rho * std::cos(ang), ly = rho * std::sin(ang);
            v[k] = {p.x + R[0][0] * lx + R[0][1] * ly, p.y + R[1][0] * lx + R[1][1] * ly,
                    p.z + R[2][0] * lx + R[2][1] * ly};
        }
        emit(out, v[0], v[1], v[2]);
    }
    return out;
}

}  // namespace rt

extern "C" int rt_scene_generate(uint32_t kind, uint64_t seed, uint32_t count, uint32_t param,
                                 float** tris, uint32_t* n_tris) {
    if (!tris || !n_tris) return rt::set_error(RT_ERR_ARG, "rt_scene_generate: null output");
    std::vector<float> v;
    try {
        if (kind == RT_SCENE_SPHERES)
            v = rt::scene_spheres(seed, count, param ? param : 10);
        else if (kind == RT_SCENE_INTERIOR)
            v = rt::scene_interior(seed, count, param);
        else
            return rt::set_error(RT_ERR_ARG, "rt_scene_generate: unknown scene kind");
    } catch (const std::bad_alloc&) {
        return rt::set_error(RT_ERR_NOMEM, "rt_scene_generate: out of host memory");
    }
    float* p = (float*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(float));
    if (!p) return rt::set_error(RT_ERR_NOMEM, "rt_scene_generate: out of host memory");
    std::memcpy(p, v.data(), v.size() * sizeof(float));
    *tris = p;
    *n_tris = (uint32_t)(v.size() / 9);
    return RT_OK;
}

extern "C" void rt_free(void* p) { std::free(p); }
