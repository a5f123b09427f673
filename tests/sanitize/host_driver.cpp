// Host-code sanitizer driver (built with -fsanitize=address,undefined by
// tests/test_host_sanitizers.py): the builder, the scene generator, the host
// ray functions and the oracle on generated scenes and triangle soups with
// repeated vertices, axis-plane and duplicated triangles, at N = 2..16,
// L = 1..16, gang 4 / 8, every builder option, on 1 and 8 threads, plus one
// build large enough for the chunked parallel sweeps.  An argument selects
// the shorter case list.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "rt.h"
#include "cpu_ref.h"

static void check(int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "%s failed: %d %s\n", what, rc, rt_last_error());
        std::exit(1);
    }
}

static std::vector<float> soup(std::mt19937& g, int n, float scale) {
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    std::vector<float> t((size_t)n * 9);
    for (int i = 0; i < n; ++i) {
        float c[3] = {u(g) * scale, u(g) * scale, u(g) * scale};
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) t[(size_t)i * 9 + v * 3 + a] = c[a] + u(g) * scale * 0.05f;
        const int kind = i % 7;
        if (kind == 1) std::memcpy(&t[(size_t)i * 9 + 3], &t[(size_t)i * 9], 12);      // repeated vertex
        if (kind == 2) std::memcpy(&t[(size_t)i * 9 + 6], &t[(size_t)i * 9 + 3], 12);  // repeated vertex
        if (kind == 3)
            for (int v = 0; v < 3; ++v) t[(size_t)i * 9 + v * 3 + 1] = t[(size_t)i * 9 + 1];  // y plane
        if (kind == 4 && i > 0) std::memcpy(&t[(size_t)i * 9], &t[(size_t)(i - 1) * 9], 36);  // duplicate
    }
    return t;
}

static void run_case(const std::vector<float>& tris, const rt_build_params& p, int idx) {
    const uint32_t n = (uint32_t)(tris.size() / 9);
    rt_host_bvh_t h;
    check(rt_build(tris.data(), n, &p, &h), "rt_build");
    rt_bvh_desc d;
    check(rt_host_bvh_desc(h, &d), "rt_host_bvh_desc");
    // touch every array
    double sum = 0;
    for (size_t i = 0; i < (size_t)d.node_count * 6 * d.node_memory; ++i)
        if (!std::isnan(d.bounds[i])) sum += d.bounds[i];
    for (uint32_t i = 0; i < d.node_count; ++i) sum += d.base[i] + d.child_type[i];
    for (size_t i = 0; i < (size_t)d.node_count * 3 * d.node_memory; ++i) sum += d.order[i];
    for (uint64_t i = 0; i < d.tri_float_count; ++i)
        if (!std::isnan(d.tri_points[i])) sum += d.tri_points[i];
    const uint64_t nl = d.tri_float_count / (9ull * d.leaf_memory);
    for (uint64_t i = 0; i < nl * d.leaf_memory; ++i) sum += d.tri_ids[i];
    orc_bvh b{d.node_memory, d.leaf_memory, d.gang, d.node_count, d.bounds, d.base, d.order,
              d.child_type, d.tri_float_count, d.tri_points};
    rt_camera cam{{0.f, 0.5f, 3.f}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, 0.866f};
    const uint32_t W = 64, H = 48, R = W * H;
    std::vector<rt_ray> rays(R);
    check(rt_camera_rays(&cam, W, H, 16, rays.data()), "rt_camera_rays");
    std::vector<orc_hit> oh(R);
    std::vector<orc_counts> oc(R);
    std::vector<uint32_t> diag(R);
    std::vector<uint8_t> occ(R);
    const orc_ray* orays = reinterpret_cast<const orc_ray*>(rays.data());
    for (int mode : {ORC_CLOSEST, ORC_SAVE_DIST, ORC_ANY})
        for (int stream : {0, 1}) {
            if (stream && mode == ORC_SAVE_DIST) continue;
            orc_trace(&b, orays, R, mode, stream, 256, 2, oh.data(), occ.data(), oc.data(), diag.data());
        }
    orc_trace(&b, orays, R, ORC_CLOSEST, 0, 256, 2, oh.data(), occ.data(), oc.data(), diag.data());
    // the emulated wide traversal on one tile
    {
        std::vector<orc_ray> tile(orays, orays + 256);
        std::vector<uint32_t> li(256);
        std::vector<int32_t> ti(256);
        orc_intersect_wide(&b, tile.data(), 256, ORC_WIDE_AXIS_BUG | ORC_WIDE_SHARED_DIST, li.data(), ti.data(),
                           oc.data(), diag.data());
    }
    std::vector<uint32_t> ids((size_t)R * 64), lens(R);
    orc_visit_sequences(&b, orays, R, ORC_CLOSEST, ids.data(), 64, lens.data());
    // hits in the ABI record, surfaces, secondary rays
    std::vector<rt_hit> hits(R);
    for (uint32_t i = 0; i < R; ++i) {
        hits[i].slot = oh[i].slot;
        hits[i].leaf = oh[i].leaf;
        hits[i].t = oh[i].t;
        hits[i].b1 = oh[i].b1;
        hits[i].b2 = oh[i].b2;
        hits[i].tri = oh[i].slot >= 0 ? d.tri_ids[(oh[i].leaf / (9 * d.leaf_memory)) * d.leaf_memory + oh[i].slot] : -1;
    }
    std::vector<float> nrm(3 * R), pos(3 * R);
    check(rt_hit_surface(h, rays.data(), hits.data(), R, nrm.data(), pos.data()), "rt_hit_surface");
    std::vector<rt_ray> sec(R);
    std::vector<uint32_t> src(R);
    uint32_t ns = 0;
    const float light[3] = {0.f, 2.f, 0.f};
    check(rt_spawn_shadow(nrm.data(), pos.data(), hits.data(), R, light, 1.f, sec.data(), src.data(), &ns),
          "rt_spawn_shadow");
    if (ns) orc_trace(&b, reinterpret_cast<const orc_ray*>(sec.data()), ns, ORC_ANY, 0, 256, 2, nullptr, occ.data(),
                      oc.data(), diag.data());
    check(rt_spawn_ambient(nrm.data(), pos.data(), hits.data(), R, 3, 0.5f, sec.data(), src.data(), &ns),
          "rt_spawn_ambient");
    uint32_t nh = 0;
    for (auto& x : hits) nh += x.slot >= 0;
    if (nh) {
        std::vector<rt_ray> diff(5000);
        check(rt_spawn_diffuse(rays.data(), nrm.data(), pos.data(), hits.data(), R, 7, 5000, 222222.f, diff.data()),
              "rt_spawn_diffuse");
        std::vector<orc_hit> dh(5000);
        std::vector<orc_counts> dc(5000);
        std::vector<uint32_t> dd(5000);
        std::vector<uint8_t> docc(5000);
        orc_trace(&b, reinterpret_cast<const orc_ray*>(diff.data()), 5000, ORC_CLOSEST, 0, 256, 2, dh.data(),
                  docc.data(), dc.data(), dd.data());
    }
    std::printf("case %d: %u tris N%u L%u G%u threads %u: %u nodes, %u hits, checksum %.3g\n", idx, n, p.branching,
                p.leaf_size, p.gang, p.threads, d.node_count, nh, sum);
    check(rt_host_bvh_free(h), "rt_host_bvh_free");
}

int main(int argc, char** argv) {
    const bool quick = argc > 1;
    std::mt19937 g(11);
    int idx = 0;
    const uint32_t Ns[] = {2, 4, 5, 8, 12, 16}, Ls[] = {1, 3, 4, 8, 16};
    for (int round = 0; round < 2; ++round)
        for (uint32_t N : Ns)
            for (uint32_t L : Ls) {
                if (quick && (idx % 3) != 0) { ++idx; continue; }
                const uint32_t G = (idx % 3 == 0) ? 8 : 4;
                std::vector<float> tris;
                if (idx % 4 == 0) {
                    float* out = nullptr;
                    uint32_t cnt = 0;
                    check(rt_scene_generate(idx % 8 == 0 ? 1 : 2, idx, 3000 + 500 * (idx % 5), 0, &out, &cnt),
                          "rt_scene_generate");
                    tris.assign(out, out + (size_t)cnt * 9);
                    rt_free(out);
                } else {
                    tris = soup(g, 200 + 300 * (idx % 7), idx % 5 == 0 ? 1e6f : 1.f);
                }
                rt_build_params p{};
                p.branching = N;
                p.leaf_size = L;
                p.gang = G;
                p.sort_each_split = (idx % 5) != 1;
                p.leaf_split_option = idx % 3;
                p.sah_factor = rt_sah_factor(L, N, G, nullptr);
                p.threads = round == 0 ? 1 : 8;
                run_case(tris, p, idx++);
            }
    // one large build to reach the chunked (parallel) sweeps and sort
    {
        float* out = nullptr;
        uint32_t cnt = 0;
        check(rt_scene_generate(2, 99, quick ? 140000 : 300000, 0, &out, &cnt), "rt_scene_generate");
        std::vector<float> tris(out, out + (size_t)cnt * 9);
        rt_free(out);
        rt_build_params p{};
        p.branching = 8;
        p.leaf_size = 8;
        p.gang = 4;
        p.sort_each_split = 1;
        p.leaf_split_option = 1;
        p.sah_factor = rt_sah_factor(8, 8, 4, nullptr);
        p.threads = 8;
        run_case(tris, p, idx++);
    }
    std::printf("all %d cases clean\n", idx);
    return 0;
}
