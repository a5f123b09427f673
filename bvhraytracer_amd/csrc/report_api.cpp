// C entry points of the renderer layer's report files and config.txt parsing
// (include/rt_report.h).  Each wraps the one C++ implementation in
// renderer.cpp; exceptions become RT_ERR_ARG with rt_last_error() set.
#include "../../include/rt_report.h"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/rt_renderer.hpp"
#include "host_common.h"

namespace {

template <typename F>
int guarded(const char* what, F&& f) {
    try {
        f();
        return RT_OK;
    } catch (const std::exception& e) {
        return rt::set_error(RT_ERR_ARG, std::string(what) + ": " + e.what());
    }
}

}  // namespace

extern "C" int rt_report_read_config(const char* path, char* text, size_t cap, size_t* len) {
    if (!path || !len || (cap && !text)) return rt::set_error(RT_ERR_ARG, "rt_report_read_config: null argument");
    std::string out;
    int rc = guarded("rt_report_read_config", [&] {
        rt::RenderConfig c;
        if (!rt::read_config(path, c)) throw rt::Error(std::string("cannot open ") + path);
        out = rt::config_text(c);
    });
    if (rc != RT_OK) return rc;
    *len = out.size();
    if (out.size() + 1 > cap)
        return rt::set_error(RT_ERR_ARG, "rt_report_read_config: text of " + std::to_string(out.size()) +
                                             " bytes does not fit the buffer");
    std::memcpy(text, out.c_str(), out.size() + 1);
    return RT_OK;
}

extern "C" int rt_report_write_perf(const char* path, const char* name, uint32_t branching, uint32_t leaf_size,
                                    uint32_t gang, uint32_t node_memory, uint32_t leaf_memory, float total_s,
                                    float ray_sum_s, float triangle_sum_s, float average_depth, uint64_t triangles) {
    if (!path || !name) return rt::set_error(RT_ERR_ARG, "rt_report_write_perf: null argument");
    return guarded("rt_report_write_perf", [&] {
        rt::write_perf(path, name, branching, leaf_size, gang, node_memory, leaf_memory, total_s, ray_sum_s,
                       triangle_sum_s, average_depth, triangles);
    });
}

extern "C" int rt_report_write_info(const char* path, const char* name, uint32_t branching, uint32_t leaf_size,
                                    uint32_t bvh_depth, const rt_trace_hist* primary,
                                    const rt_trace_hist* secondary) {
    if (!path || !name || !primary || !secondary)
        return rt::set_error(RT_ERR_ARG, "rt_report_write_info: null argument");
    if (bvh_depth > RT_HIST_DEPTHS)
        return rt::set_error(RT_ERR_ARG, "rt_report_write_info: bvh_depth above the histogram's depth bins");
    return guarded("rt_report_write_info",
                   [&] { rt::write_info(path, name, branching, leaf_size, bvh_depth, *primary, *secondary); });
}

extern "C" int rt_report_write_workgroup_analysis(const char* path, const char* name, const char* problem,
                                                  int32_t camera_id, uint32_t width, uint32_t height, uint32_t wg,
                                                  const rt_step_stat* primary, const uint32_t* primary_len,
                                                  const rt_step_stat* secondary, const uint32_t* secondary_len,
                                                  const uint32_t* primary_nodes, const uint32_t* primary_leaves,
                                                  const uint32_t* secondary_nodes, const uint32_t* secondary_leaves,
                                                  int32_t wide_alternative, char* paths, size_t cap) {
    if (!path || !name || !problem || !primary_len || !secondary_len || !primary_nodes || !primary_leaves ||
        !secondary_nodes || !secondary_leaves || (cap && !paths))
        return rt::set_error(RT_ERR_ARG, "rt_report_write_workgroup_analysis: null argument");
    if (wg == 0 || width % wg != 0 || height < wg)
        return rt::set_error(RT_ERR_ARG, "rt_report_write_workgroup_analysis: no full tile of the workgroup size");
    const size_t tiles = (size_t)(width / wg) * (height / wg), wgsq = (size_t)wg * wg;
    std::vector<std::vector<rt_step_stat>> p(tiles), s(tiles);
    size_t op = 0, os = 0;
    for (size_t i = 0; i < tiles; ++i) {
        if (primary_len[i] && !primary) return rt::set_error(RT_ERR_ARG, "rt_report_write_workgroup_analysis: null records");
        if (secondary_len[i] && !secondary)
            return rt::set_error(RT_ERR_ARG, "rt_report_write_workgroup_analysis: null records");
        p[i].assign(primary + op, primary + op + primary_len[i]);
        s[i].assign(secondary + os, secondary + os + secondary_len[i]);
        op += primary_len[i];
        os += secondary_len[i];
    }
    rt::TileVisits pv, sv;
    pv.nodes.assign(primary_nodes, primary_nodes + tiles * wgsq);
    pv.leaves.assign(primary_leaves, primary_leaves + tiles * wgsq);
    sv.nodes.assign(secondary_nodes, secondary_nodes + tiles * wgsq);
    sv.leaves.assign(secondary_leaves, secondary_leaves + tiles * wgsq);
    std::string list;
    const int rc = guarded("rt_report_write_workgroup_analysis", [&] {
        for (const std::string& f : rt::write_workgroup_analysis(path, name, problem, camera_id, width, height, wg,
                                                                 p, s, pv, sv, wide_alternative != 0))
            list += f + "\n";
    });
    if (rc != RT_OK) return rc;
    if (cap) {
        if (list.size() + 1 > cap)
            return rt::set_error(RT_ERR_ARG, "rt_report_write_workgroup_analysis: paths do not fit the buffer");
        std::memcpy(paths, list.c_str(), list.size() + 1);
    }
    return RT_OK;
}

extern "C" int rt_report_tile_order_to_rows(const uint8_t* tiles, uint32_t width, uint32_t height, uint32_t wg,
                                            uint32_t channels, uint8_t* rows) {
    const size_t bytes = (size_t)width * height * channels;
    if (bytes && (!tiles || !rows)) return rt::set_error(RT_ERR_ARG, "rt_report_tile_order_to_rows: null argument");
    return guarded("rt_report_tile_order_to_rows", [&] {
        const std::vector<uint8_t> out =
            rt::tile_order_to_rows(std::vector<uint8_t>(tiles, tiles + bytes), width, height, wg, channels);
        if (bytes) std::memcpy(rows, out.data(), bytes);
    });
}

extern "C" int rt_report_write_png(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height) {
    const size_t bytes = (size_t)width * height * 4;
    if (!path || (bytes && !rgba)) return rt::set_error(RT_ERR_ARG, "rt_report_write_png: null argument");
    return guarded("rt_report_write_png",
                   [&] { rt::write_png(path, std::vector<uint8_t>(rgba, rgba + bytes), width, height); });
}

extern "C" int rt_report_ambient_gray(const uint8_t* hit, const uint32_t* occluded, uint32_t n, uint32_t samples,
                                      uint8_t* gray) {
    if (n && (!hit || !occluded || !gray)) return rt::set_error(RT_ERR_ARG, "rt_report_ambient_gray: null argument");
    for (uint32_t i = 0; i < n; ++i) gray[i] = rt::ambient_gray(hit[i] != 0, occluded[i], samples);
    return RT_OK;
}

extern "C" int rt_report_ambient_distance(const float* tris, uint32_t n_tris, float* out) {
    if (!out || (n_tris && !tris)) return rt::set_error(RT_ERR_ARG, "rt_report_ambient_distance: null argument");
    *out = rt::ambient_distance(std::vector<float>(tris, tris + (size_t)n_tris * 9));
    return RT_OK;
}
