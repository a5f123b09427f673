// Sanitizer driver for the host-only parts of the C++ renderer layer
// (config parsing, sweep, SAH factors, _Info.txt and workgroup files, tile
// order, PNG, scenario geometry) and their C entry points (rt_report.h);
// argv[1] is the output directory.
#include <cstdio>
#include <fstream>
#include <random>

#include "rt_renderer.hpp"
#include "rt_report.h"

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
    // the C entry points over the same code: a config text buffer too small
    // first, flattened step records, null arguments refused
    {
        size_t len = 0;
        char small[8];
        if (rt_report_read_config((dir + "/config.txt").c_str(), small, sizeof small, &len) != RT_ERR_ARG) return 3;
        std::vector<char> text(len + 1);
        if (rt_report_read_config((dir + "/config.txt").c_str(), text.data(), text.size(), &len) != RT_OK) return 4;
        if (rt_report_read_config((dir + "/missing.txt").c_str(), text.data(), text.size(), &len) != RT_ERR_ARG)
            return 5;
        if (rt_report_write_perf((dir + "/x_Perf.txt").c_str(), "scene", 8, 8, 4, 8, 8, 1.5f, 1.0f, 0.25f, 9.5f,
                                 100002) != RT_OK)
            return 6;
        rt_trace_hist a{};
        a.rays = 10; a.node_visits = 300; a.leaf_visits = 70; a.aabb_tests = 2000; a.aabb_hits = 400;
        a.tri_tests = 200; a.tri_hits = 9;
        a.node_depth[3] = 5; a.child_fullness[4] = 2; a.leaf_fullness[2] = 3;
        if (rt_report_write_info((dir + "/c_Info.txt").c_str(), "scene", 8, 8, 12, &a, &a) != RT_OK) return 7;
        if (rt_report_write_info((dir + "/c_Info.txt").c_str(), "scene", 8, 8, 65, &a, &a) != RT_ERR_ARG) return 8;
        const unsigned wg = 4, W = 8, H = 4, tiles = 2, wgsq = wg * wg;
        std::vector<rt_step_stat> recs;
        std::vector<uint32_t> lens;
        for (unsigned t = 0; t < tiles; ++t) {
            rt_step_stat r0{}, r1{};
            r0.node_work = wgsq;
            r0.node_unique = 1;
            r1.leaf_work = wgsq;
            r1.leaf_unique = 2;
            r1.terminations = wgsq;
            recs.push_back(r0);
            recs.push_back(r1);
            lens.push_back(2);
        }
        std::vector<uint32_t> visits(tiles * wgsq, 1);
        char paths[4096];
        if (rt_report_write_workgroup_analysis(dir.c_str(), "scene", "prob", 1, W, H, wg, recs.data(), lens.data(),
                                               recs.data(), lens.data(), visits.data(), visits.data(), visits.data(),
                                               visits.data(), 1, paths, sizeof paths) != RT_OK)
            return 9;
        if (rt_report_write_workgroup_analysis(dir.c_str(), "scene", "prob", 1, W, H, 3, recs.data(), lens.data(),
                                               recs.data(), lens.data(), visits.data(), visits.data(), visits.data(),
                                               visits.data(), 1, paths, sizeof paths) != RT_ERR_ARG)
            return 10;
        std::vector<uint8_t> px(W * H * 4), rows(W * H * 4);
        for (auto& x : px) x = (uint8_t)g();
        if (rt_report_tile_order_to_rows(px.data(), W, H, wg, 4, rows.data()) != RT_OK) return 11;
        if (rt_report_write_png((dir + "/c.png").c_str(), rows.data(), W, H) != RT_OK) return 12;
        std::vector<uint8_t> hit(W * H, 1), gray(W * H);
        std::vector<uint32_t> occ(W * H, 1);
        if (rt_report_ambient_gray(hit.data(), occ.data(), W * H, 3, gray.data()) != RT_OK) return 13;
        float dist = 0;
        if (rt_report_ambient_distance(tris.data(), (uint32_t)(tris.size() / 9), &dist) != RT_OK) return 14;
        if (rt_report_write_png(nullptr, rows.data(), W, H) != RT_ERR_ARG) return 15;
        std::printf("C entry points: config %zu bytes, ambient distance %g\n", len, dist);
    }
    std::printf("renderer host parts clean\n");
    return 0;
}
