// Sanitizer driver for the host-only parts of the C++ renderer layer
// (config parsing, sweep, SAH factors, _Info.txt and workgroup files, tile
// order, PNG, scenario geometry); argv[1] is the output directory.
#include <cstdio>
#include <fstream>
#include <random>

#include "rt_renderer.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::mt19937 g(3);
    const std::string dir = argv[1];
    // config parsing: odd lines, substring keys, bad numbers are reported by
    // std::stoi in the reference too, so only well-formed numbers here
    {
        std::ofstream f(dir + "/config.txt");
        f << "# comment\n\nminLeafSize = 4\nmaxLeafSize=16\nleafStep = 4\nminBranch =2\nmaxBranch = 16\n"
             "branchStep = 2\nxRes = 640\nyRes = 480\nworkGroupSize = 8\nsaveImage = true\nscenarios = 0\n"
             "sortEachSplit false\nwideRender = true  # trailing\nambientSampleCount = 3\n"
             "doNodeMemoryTest = true\ndoLeafMemoryTest = true\nleafSplitOption = 2\n";
    }
    rt::RenderConfig cfg;
    if (!rt::read_config(dir + "/config.txt", cfg)) return 1;
    rt::sanity_clamp(cfg);
    auto sw = rt::sweep(cfg);
    std::printf("sweep %zu pairs, warnings %zu\n", sw.size(), cfg.warnings.size());
    for (auto& p : sw) (void)rt::load_sah_factor(p.first, p.second, 4);
    // _Info.txt from synthetic counters
    for (unsigned N : {2u, 4u, 8u, 16u}) {
        rt_trace_hist a{}, b{};
        a.rays = b.rays = 1000;
        a.node_visits = 30000; a.leaf_visits = 7000; a.aabb_tests = 200000; a.aabb_hits = 40000;
        a.tri_tests = 20000; a.tri_hits = 900;
        b = a;
        for (int d = 0; d < RT_HIST_DEPTHS; ++d) { a.node_depth[d] = g() % 1000; a.leaf_depth[d] = g() % 500; }
        for (int k = 0; k < RT_HIST_FULL; ++k) { a.child_fullness[k] = g() % 100; a.leaf_fullness[k] = g() % 100; }
        rt::write_info(dir + "/scene_Info.txt", "scene", N, 4, 12, a, b);
    }
    // workgroup files from random step records
    for (unsigned wg : {4u, 8u, 16u}) {
        const unsigned W = 4 * wg, H = 3 * wg, tiles = (W / wg) * (H / wg);
        std::vector<std::vector<rt_step_stat>> pri(tiles), sec(tiles);
        rt::TileVisits pv, sv;
        for (auto* v : {&pri, &sec})
            for (auto& t : *v) {
                // consistent records: every active ray visits one entry per
                // step, the last step ends the rays still active
                const unsigned steps = 2 + g() % 40;
                unsigned active = wg * wg;
                t.resize(steps);
                for (unsigned k = 0; k < steps; ++k) {
                    rt_step_stat r{};
                    if (k == 0) { r.terminations = g() % 2; active -= r.terminations; t[k] = r; continue; }
                    r.node_work = active ? (unsigned)(g() % (active + 1)) : 0;
                    r.leaf_work = active - r.node_work;
                    r.node_unique = r.node_work ? 1 + g() % r.node_work : 0;
                    r.leaf_unique = r.leaf_work ? 1 + g() % r.leaf_work : 0;
                    r.terminations = (k == steps - 1) ? active : (unsigned)(g() % (active / 4 + 1));
                    active -= r.terminations;
                    t[k] = r;
                }
            }
        pv.nodes.resize(W * H); pv.leaves.resize(W * H);
        sv.nodes.resize(W * H); sv.leaves.resize(W * H);
        for (auto& x : pv.nodes) x = g() % 60;
        for (auto& x : pv.leaves) x = g() % 20;
        for (auto& x : sv.nodes) x = g() % 60;
        for (auto& x : sv.leaves) x = g() % 20;
        auto files = rt::write_workgroup_analysis(dir, "scene", "prob", 0, W, H, wg, pri, sec, pv, sv, wg % 2 == 0);
        std::printf("wg %u: %zu files\n", wg, files.size());
    }
    // image paths: tile order with a partial last tile row, PNG
    for (unsigned h : {48u, 50u, 61u}) {
        const unsigned w = 64, wg = 16;
        std::vector<uint8_t> tiles((size_t)w * h * 4);
        for (auto& x : tiles) x = (uint8_t)g();
        auto rows = rt::tile_order_to_rows(tiles, w, h, wg, 4);
        rt::write_png(dir + "/img.png", rows, w, h);
    }
    for (unsigned s : {0u, 1u, 7u}) (void)rt::ambient_gray(true, s / 2, s);
    std::string name;
    rt::Vec3 cp, ct;
    auto tris = rt::scenario_triangles(0, name, cp, ct);
    std::printf("scenario 0: %zu floats, ambient distance %g\n", tris.size(), rt::ambient_distance(tris));
    std::printf("renderer host parts clean\n");
    return 0;
}
