// C++ host layer (rt_renderer.hpp): config.txt, the N/L sweep, RayTracer,
// CameraFast and the _Perf.txt / _Perf.png reports over the C ABI.
//
//   read_config / sanity_clamp / sweep   src/rayTracer.cpp:97-122,568-576,756-877
//   FastNodeManager                      src/accelerationStructure/fastNodeManager.h:1232-1763
//   CameraFast::renderImages / renderImage / the PNG   src/cameraFast.h:49-582,829-852
//   RayTracer::renderImage               src/rayTracer.cpp:588-745 (performance branch)
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>

#include "../../include/rt_renderer.hpp"

namespace rt {

namespace {

[[noreturn]] void fail(const std::string& what) { throw Error(what + ": " + rt_last_error()); }

void check(int rc, const char* what) {
    if (rc != RT_OK) fail(what);
}

bool has(const std::string& line, const char* key) { return line.find(key) != std::string::npos; }

// std::stoi of the text after the line's first '=' (the whole line when it
// has none: find() + 1 wraps npos to 0)
int value_after_eq(const std::string& line) {
    const std::string s = line.substr(line.find('=') + 1);
    try {
        return std::stoi(s);
    } catch (const std::exception&) {
        throw Error("config.txt: no integer in '" + line + "'");
    }
}

std::vector<int> int_list(const std::string& line) {
    std::vector<int> out;
    std::string s = line.substr(line.find('=') + 1);
    size_t pos;
    while ((pos = s.find(',')) != std::string::npos) {
        out.push_back(value_after_eq("=" + s.substr(0, pos)));
        s.erase(0, pos + 1);
    }
    out.push_back(value_after_eq("=" + s));
    return out;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

bool read_config(const std::string& path, RenderConfig& c) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    const std::pair<const char*, int*> ints[] = {
        {"minLeafSize", &c.minLeafSize},   {"maxLeafSize", &c.maxLeafSize},
        {"leafStep", &c.leafStep},         {"minBranch", &c.minBranch},
        {"maxBranch", &c.maxBranch},       {"branchStep", &c.branchStep},
        {"subdivisionStart", &c.subdivisionStart}, {"subdivisionEnd", &c.subdivisionEnd},
        {"subdivisionStep", &c.subdivisionStep},   {"renderType", &c.renderType},
        {"ambientSampleCount", &c.ambientSampleCount}, {"workGroupSize", &c.workGroupSize},
        {"leafSplitOption", &c.leafSplitOption},   {"xRes", &c.xRes},
        {"yRes", &c.yRes}};
    const std::pair<const char*, bool*> bools[] = {
        {"saveImage", &c.saveImage},       {"saveDepthDetailedImage", &c.saveDepthDetailedImage},
        {"bvhAnalysis", &c.bvhAnalysis},   {"saveBvhImage", &c.saveBvhImage},
        {"sortEachSplit", &c.sortEachSplit}, {"renderAnalysisImage", &c.renderAnalysisImage},
        {"castShadows", &c.castShadows},   {"doPerformanceTest", &c.doPerformanceTest},
        {"doNodeMemoryTest", &c.doNodeMemoryTest}, {"doLeafMemoryTest", &c.doLeafMemoryTest},
        {"saveRayTimes", &c.saveRayTimes}, {"doWorkGroupAnalysis", &c.doWorkGroupAnalysis},
        {"saveDistance", &c.saveDistance}, {"wideRender", &c.wideRender},
        {"wideAlternative", &c.wideAlternative}, {"renderAllOptions", &c.renderAllOptions},
        {"doCacheAnalysis", &c.doCacheAnalysis}};
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        for (const auto& kv : ints)
            if (has(line, kv.first)) *kv.second = value_after_eq(line);
        for (const auto& kv : bools) {
            if (!has(line, kv.first)) continue;
            if (has(line, "true")) {
                *kv.second = true;
            } else if (has(line, "false")) {
                *kv.second = false;
            } else {
                c.warnings.push_back(std::string(kv.first) + "value written wrong -> default = false");
                *kv.second = false;
            }
        }
        if (has(line, "scenario"))
            for (int v : int_list(line)) c.scenarios.push_back(v);
        if (c.doCacheAnalysis && has(line, "cacheSize")) {
            // the reference's loop does not advance past a size it rejects
            // (rayTracer.cpp:843-849); here such a size is skipped
            for (int v : int_list(line)) {
                if (v % 8 == 0) c.cacheSizes.push_back(v);
                else c.warnings.push_back("cachesize must be divideable by 8: " + std::to_string(v) + " is ignored");
            }
            if (c.cacheSizes.empty()) c.cacheSizes.push_back(256);
        }
    }
    return true;
}

std::string config_text(const RenderConfig& c) {
    const std::pair<const char*, int> ints[] = {
        {"minLeafSize", c.minLeafSize},     {"maxLeafSize", c.maxLeafSize},   {"leafStep", c.leafStep},
        {"minBranch", c.minBranch},         {"maxBranch", c.maxBranch},       {"branchStep", c.branchStep},
        {"subdivisionStart", c.subdivisionStart}, {"subdivisionEnd", c.subdivisionEnd},
        {"subdivisionStep", c.subdivisionStep},   {"renderType", c.renderType},
        {"ambientSampleCount", c.ambientSampleCount}, {"workGroupSize", c.workGroupSize},
        {"leafSplitOption", c.leafSplitOption},   {"xRes", c.xRes},             {"yRes", c.yRes}};
    const std::pair<const char*, bool> bools[] = {
        {"saveImage", c.saveImage},         {"saveDepthDetailedImage", c.saveDepthDetailedImage},
        {"bvhAnalysis", c.bvhAnalysis},     {"saveBvhImage", c.saveBvhImage},
        {"sortEachSplit", c.sortEachSplit}, {"renderAnalysisImage", c.renderAnalysisImage},
        {"castShadows", c.castShadows},     {"doPerformanceTest", c.doPerformanceTest},
        {"doNodeMemoryTest", c.doNodeMemoryTest}, {"doLeafMemoryTest", c.doLeafMemoryTest},
        {"saveRayTimes", c.saveRayTimes},   {"doWorkGroupAnalysis", c.doWorkGroupAnalysis},
        {"saveDistance", c.saveDistance},   {"wideRender", c.wideRender},
        {"wideAlternative", c.wideAlternative}, {"renderAllOptions", c.renderAllOptions},
        {"doCacheAnalysis", c.doCacheAnalysis}};
    std::string out;
    for (const auto& kv : ints) out += std::string(kv.first) + "=" + std::to_string(kv.second) + "\n";
    for (const auto& kv : bools) out += std::string(kv.first) + "=" + (kv.second ? "true" : "false") + "\n";
    auto list = [&out](const char* k, const std::vector<int>& v) {
        out += std::string(k) + "=";
        for (size_t i = 0; i < v.size(); ++i) out += (i ? "," : "") + std::to_string(v[i]);
        out += "\n";
    };
    list("scenarios", c.scenarios);
    list("cacheSizes", c.cacheSizes);
    for (const auto& w : c.warnings) out += "warning=" + w + "\n";
    return out;
}

void sanity_clamp(RenderConfig& c) {
    c.maxBranch = std::max(c.maxBranch, c.minBranch);
    c.minBranch = std::min(c.maxBranch, c.minBranch);
    c.maxLeafSize = std::max(c.maxLeafSize, c.minLeafSize);
    c.minLeafSize = std::min(c.maxLeafSize, c.minLeafSize);
}

std::vector<std::pair<int, int>> sweep(const RenderConfig& c) {
    std::vector<std::pair<int, int>> out;
    if (c.leafStep <= 0 || c.branchStep <= 0) throw Error("config.txt: leafStep and branchStep must be positive");
    for (int l = c.minLeafSize; l <= c.maxLeafSize; l += c.leafStep)
        for (int b = c.minBranch; b <= c.maxBranch; b += c.branchStep) out.emplace_back(l, b);
    return out;
}

float load_sah_factor(int leafSize, int branchingFactor, int gangSize, const std::string& table) {
    return rt_sah_factor((uint32_t)leafSize, (uint32_t)branchingFactor, (uint32_t)gangSize,
                         table.empty() ? nullptr : table.c_str());
}

// ---------------------------------------------------------------------------

FastNodeManager::FastNodeManager(const std::vector<float>& tris, unsigned branching, unsigned leaf, unsigned gang,
                                 float sahFactor, bool sortEachSplit, int leafSplitOption, unsigned threads,
                                 unsigned nodeMemory_, unsigned leafMemory_) {
    if (tris.size() % 9 != 0) throw Error("FastNodeManager: triangle array is not a multiple of 9 floats");
    rt_build_params p{branching, leaf, gang, sortEachSplit ? 1u : 0u, leafSplitOption, sahFactor, threads,
                      nodeMemory_, leafMemory_};
    check(rt_build(tris.data(), (uint32_t)(tris.size() / 9), &p, &host_), "rt_build");
    if (rt_host_bvh_desc(host_, &desc_) != RT_OK) {
        rt_host_bvh_free(host_);
        fail("rt_host_bvh_desc");
    }
    if (rt_upload_bvh(&desc_, &scene_) != RT_OK) {
        rt_host_bvh_free(host_);
        fail("rt_upload_bvh");
    }
    branchingFactor = desc_.branching;
    leafSize = desc_.leaf_size;
    gangSize = desc_.gang;
    nodeMemory = desc_.node_memory;
    leafMemory = desc_.leaf_memory;
    averageBvhDepth = desc_.average_depth;
    triangleCount = desc_.triangle_count;
}

FastNodeManager::~FastNodeManager() {
    if (scene_) rt_release(scene_);
    if (host_) rt_host_bvh_free(host_);
}

void FastNodeManager::intersect(const std::vector<rt_ray>& rays, std::vector<rt_hit>& hits,
                                std::vector<rt_counts>* counts, bool saveDistance, bool wide) {
    hits.resize(rays.size());
    uint32_t flags = saveDistance ? RT_SAVE_DIST : RT_CLOSEST;
    if (wide) flags |= RT_STREAM;
    if (counts) {
        counts->resize(rays.size());
        flags |= RT_COUNTS;
    }
    check(rt_trace(scene_, rays.data(), (uint32_t)rays.size(), flags, hits.data(), counts ? counts->data() : nullptr,
                   nullptr),
          "rt_trace");
}

void FastNodeManager::intersectSecondary(const std::vector<rt_ray>& rays, std::vector<uint8_t>& occluded,
                                         std::vector<rt_counts>* counts, bool wide) {
    occluded.resize(rays.size());
    uint32_t flags = RT_ANY | (wide ? RT_STREAM : 0u);
    if (counts) {
        counts->resize(rays.size());
        flags |= RT_COUNTS;
    }
    check(rt_trace(scene_, rays.data(), (uint32_t)rays.size(), flags, occluded.data(),
                   counts ? counts->data() : nullptr, nullptr),
          "rt_trace");
}

void FastNodeManager::getSurfaceNormalPosition(const std::vector<rt_ray>& rays, const std::vector<rt_hit>& hits,
                                               std::vector<float>& normals, std::vector<float>& positions) const {
    if (rays.size() != hits.size()) throw Error("getSurfaceNormalPosition: rays and hits differ in length");
    normals.resize(rays.size() * 3);
    positions.resize(rays.size() * 3);
    check(rt_hit_surface(host_, rays.data(), hits.data(), (uint32_t)rays.size(), normals.data(), positions.data()),
          "rt_hit_surface");
}

void FastNodeManager::histograms(const std::vector<rt_ray>& rays, bool anyHit, rt_trace_hist& hist,
                                 std::vector<rt_hit>* hits) {
    if (hits) hits->resize(rays.size());
    if (rays.empty()) return;
    rt_trace_hist h{};
    std::vector<uint8_t> out(rays.size() * (anyHit ? 1 : sizeof(rt_hit)));
    check(rt_trace_histograms(scene_, rays.data(), (uint32_t)rays.size(), anyHit ? RT_ANY : RT_CLOSEST, out.data(), &h,
                              nullptr),
          "rt_trace_histograms");
    if (hits && !anyHit) std::memcpy(hits->data(), out.data(), out.size());
    uint64_t* dst = reinterpret_cast<uint64_t*>(&hist);
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&h);
    for (size_t i = 0; i < sizeof(rt_trace_hist) / 8; ++i) dst[i] += src[i];
}

std::vector<std::vector<rt_step_stat>> FastNodeManager::workgroupSteps(const std::vector<rt_ray>& rays, unsigned tile,
                                                                       bool anyHit, TileVisits* visits) {
    if (tile == 0 || rays.size() % tile != 0) throw Error("workgroupSteps: the ray count must be a multiple of tile");
    const size_t tiles = rays.size() / tile;
    std::vector<rt_counts> cnt(rays.size());
    std::vector<uint32_t> lens(std::max<size_t>(1, tiles));
    uint32_t cap = 256;
    std::vector<rt_step_stat> rec;
    for (;;) {
        rec.assign(std::max<size_t>(1, tiles * cap), rt_step_stat{});
        check(rt_trace_steps(scene_, rays.data(), (uint32_t)rays.size(), anyHit ? RT_ANY : RT_CLOSEST, tile, cap,
                             rec.data(), lens.data(), cnt.data(), nullptr),
              "rt_trace_steps");
        uint32_t need = 0;
        for (size_t i = 0; i < tiles; ++i) need = std::max(need, lens[i]);
        if (need <= cap) break;
        cap = need;
    }
    std::vector<std::vector<rt_step_stat>> out(tiles);
    for (size_t i = 0; i < tiles; ++i) out[i].assign(rec.begin() + i * cap, rec.begin() + i * cap + lens[i]);
    if (visits) {
        visits->nodes.resize(rays.size());
        visits->leaves.resize(rays.size());
        for (size_t r = 0; r < rays.size(); ++r) {
            visits->nodes[r] = cnt[r].nodes;
            visits->leaves[r] = cnt[r].leaves;
        }
    }
    return out;
}

// ---------------------------------------------------------------------------

namespace {

void info_block(std::ofstream& f, const char* label, const rt_trace_hist& h, unsigned branching) {
    const float factor = 1 / (float)h.rays;
    f << label << " intersections node: " << h.node_visits * factor << std::endl;
    f << label << " aabb intersections: " << h.aabb_tests * factor << std::endl;
    f << label << " aabb success ration: " << h.aabb_hits / (float)h.aabb_tests << std::endl;
    const float waste = (h.leaf_visits + h.node_visits - h.rays) / (float)(h.node_visits * branching);
    f << label << " waste factor: " << 1 - waste << std::endl;
    f << label << " intersections leaf: " << h.leaf_visits * factor << std::endl;
    f << label << " primitive intersections: " << h.tri_tests * factor << std::endl;
    f << label << " primitive success ratio: " << h.tri_hits / (float)h.tri_tests << std::endl;
}

void fullness(std::ofstream& f, const char* title, const char* avgKey, const uint64_t* counts, unsigned size,
              float factor) {
    std::vector<uint64_t> c(size, 0);
    for (unsigned i = 0; i < size && i < RT_HIST_FULL; ++i) c[i] = counts[i];
    float sum = 0;
    for (size_t i = 0; i < c.size(); ++i) sum += c[i] * i;
    int total = 0;  // std::accumulate(..., 0) sums into an int
    for (uint64_t v : c) total = (int)(total + v);
    sum /= total;
    f << title << std::endl;
    f << avgKey << std::to_string(sum) << std::endl;
    for (size_t i = 0; i < c.size(); ++i) f << i << " : " << c[i] * factor << std::endl;
}

}  // namespace

void write_info(const std::string& path, const std::string& name, unsigned branching, unsigned leafSize,
                unsigned bvhDepth, const rt_trace_hist& p, const rt_trace_hist& s) {
    std::ofstream f(path);
    if (!f.is_open()) throw Error("write_info: cannot open " + path);
    f << "scenario " << name << " with branching factor of " << branching << " and leafsize of " << leafSize
      << std::endl;
    f << "intersections counts are normalized per Ray" << std::endl << std::endl;
    info_block(f, "primary", p, branching);
    f << std::endl;
    info_block(f, "secondary", s, branching);
    f << std::endl;
    const float factor = 1 / (float)p.rays;
    fullness(f, "intersections with nodes with x children :", "average child fullness: ", p.child_fullness,
             branching + 1, factor);
    f << std::endl;
    fullness(f, "intersections with leaf nodes with x primitives :", "average leaf fullness: ", p.leaf_fullness,
             leafSize + 1, factor);
    const std::pair<const char*, const uint64_t*> sections[] = {{"primary node", p.node_depth},
                                                                {"primary leaf", p.leaf_depth},
                                                                {"secondary node", s.node_depth},
                                                                {"secondary leaf", s.leaf_depth}};
    for (int k = 0; k < 4; ++k) {
        const float fac = 1 / (float)(k < 2 ? p.rays : s.rays);
        f << std::endl;
        f << sections[k].first << " intersections at depth x :" << std::endl;
        for (unsigned i = 0; i < bvhDepth; ++i)
            f << i << " : " << (i < RT_HIST_DEPTHS ? sections[k].second[i] : 0ull) * fac << std::endl;
    }
}

namespace {

struct Whisker {
    int median, min, max;
    float avg, sd;
};

// cameraData.h:703-779
std::vector<Whisker> whisker_rows(const std::vector<std::vector<rt_step_stat>>& tiles, unsigned wgsq) {
    const int medianId = (int)(wgsq * 0.5f);
    std::vector<Whisker> rows;
    for (const auto& t : tiles) {
        std::vector<int> depth;
        int workSum = 0;
        for (size_t tId = 0; tId < t.size(); ++tId) {
            workSum += (int)tId * (int)t[tId].terminations;
            for (uint32_t j = 0; j < t[tId].terminations; ++j) depth.push_back((int)tId);
        }
        if (depth.size() < wgsq) throw Error("workgroup analysis: a tile has fewer terminations than rays");
        const float avg = workSum / (float)wgsq;
        float variance = 0;
        for (unsigned j = 0; j < wgsq; ++j) variance += std::pow(depth[j] - avg, 2);
        variance /= (int)wgsq;
        rows.push_back({depth[medianId], depth[0], depth[wgsq - 1], avg, std::sqrt(variance)});
    }
    std::stable_sort(rows.begin(), rows.end(), [](const Whisker& a, const Whisker& b) {
        return a.median == b.median ? a.max < b.max : a.median < b.median;
    });
    return rows;
}

struct StepAvg {
    float nodeWork = 0, nodeUnique = 0, leafWork = 0, leafUnique = 0, terminated = 0;
    uint32_t nuMax = 0, nuMin = 0, luMax = 0, luMin = 0;
};

// one half of a _WorkGroupData.txt row (cameraData.h:846-949)
StepAvg step_average(const std::vector<std::vector<rt_step_stat>>& tiles, size_t j, unsigned wgsq, float tilesF) {
    StepAvg a;
    int count = 0, nuCount = 0, luCount = 0;
    uint32_t nuMax = 0, nuMin = wgsq + 1, luMax = 0, luMin = wgsq + 1;
    for (const auto& t : tiles) {
        if (t.size() <= j) continue;
        ++count;
        a.nodeWork += t[j].node_work;
        a.leafWork += t[j].leaf_work;
        a.nodeUnique += t[j].node_unique;
        a.leafUnique += t[j].leaf_unique;
        a.terminated += t[j].terminations;
        if (t[j].node_unique != 0) {
            ++nuCount;
            nuMax = std::max(t[j].node_unique, nuMax);
            nuMin = std::min(t[j].node_unique, nuMin);
        }
        if (t[j].leaf_unique != 0) {
            ++luCount;
            luMax = std::max(t[j].leaf_unique, luMax);
            luMin = std::min(t[j].leaf_unique, luMin);
        }
    }
    count = std::max(count, 1);
    nuCount = std::max(nuCount, 1);
    luCount = std::max(luCount, 1);
    a.nuMax = nuMax;
    a.nuMin = nuMin == wgsq + 1 ? 0 : nuMin;
    a.luMax = luMax;
    a.luMin = luMin == wgsq + 1 ? 0 : luMin;
    a.nodeWork /= (float)count;
    a.leafWork /= (float)count;
    a.nodeUnique /= (float)nuCount;
    a.leafUnique /= (float)luCount;
    a.terminated /= tilesF;
    return a;
}

}  // namespace

std::vector<std::string> write_workgroup_analysis(const std::string& path, const std::string& name,
                                                  const std::string& problem, int cameraId, unsigned width,
                                                  unsigned height, unsigned wg,
                                                  const std::vector<std::vector<rt_step_stat>>& primary,
                                                  const std::vector<std::vector<rt_step_stat>>& secondary,
                                                  const TileVisits& pv, const TileVisits& sv, bool wideAlternative) {
    const unsigned wgsq = wg * wg;
    const size_t tiles = (size_t)(width / wg) * (height / wg);
    if (primary.size() != tiles || secondary.size() != tiles)
        throw Error("write_workgroup_analysis: expected " + std::to_string(tiles) + " tiles");
    if (pv.nodes.size() < tiles * wgsq || pv.leaves.size() < tiles * wgsq || sv.nodes.size() < tiles * wgsq ||
        sv.leaves.size() < tiles * wgsq)
        throw Error("write_workgroup_analysis: one visit count per ray of the full tiles");
    const std::string dir = path + "/WorkGroupSize_" + std::to_string(wg) + "_Version_" + (wideAlternative ? "1" : "0");
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    const std::string stem = dir + "/" + name + problem + "_c" + std::to_string(cameraId);
    const std::vector<std::string> files = {stem + "_PrimaryWorkGroupWiskerPlot.txt",
                                            stem + "_SecondaryWorkGroupWiskerPlot.txt", stem + "_WorkGroupData.txt",
                                            stem + "_WorkGroupUniqueWork.txt"};
    auto open = [](const std::string& fn) {
        std::ofstream f(fn);
        if (!f.is_open()) throw Error("write_workgroup_analysis: cannot open " + fn);
        return f;
    };
    for (int k = 0; k < 2; ++k) {
        std::ofstream f = open(files[k]);
        f << "median, min, max, lowerStdDeviation, upperStdDeviation" << std::endl;
        for (const Whisker& w : whisker_rows(k == 0 ? primary : secondary, wgsq))
            f << w.median << ", " << w.min << ", " << w.max << ", " << w.avg << ", " << w.sd << std::endl;
    }
    {
        // the reference's size loop tests the primary size twice, so the steps
        // are those of the longest primary tile
        size_t maxSize = 0;
        for (const auto& t : primary) maxSize = std::max(maxSize, t.size());
        const float tilesF = (float)(width / wg) * (height / wg);
        std::ofstream f = open(files[2]);
        f << "stepId, avgPrimaryNodeWork, avgPrimaryNodeUnique, avgPrimaryLeafWork, avgPrimaryLeafUnique, "
             "avgPrimaryRayTermination, primaryNodeWorkMax, primaryNodeWorkMin, primaryLeafWorkMax, "
             "primaryLeafWorkMin, avgSecondaryNodeWork, avgSecondaryNodeUnique, avgSecondaryLeafWork, "
             "avgSecondaryLeafUnique, avgSecondaryRayTermination, secondaryNodeWorkMax, secondaryNodeWorkMin, "
             "secondaryLeafWorkMax, secondaryLeafWorkMin"
          << std::endl;
        float cumP = 0, cumS = 0;
        for (size_t j = 0; j < maxSize; ++j) {
            StepAvg p = step_average(primary, j, wgsq, tilesF), s = step_average(secondary, j, wgsq, tilesF);
            if (j != 0) {
                p.terminated += cumP;
                s.terminated += cumS;
            }
            cumP = p.terminated;
            cumS = s.terminated;
            f << j << ", " << p.nodeWork << ", " << p.nodeUnique << ", " << p.leafWork << ", " << p.leafUnique << ", "
              << p.terminated << ", " << p.nuMax << ", " << p.nuMin << ", " << p.luMax << ", " << p.luMin << ", "
              << s.nodeWork << ", " << s.nodeUnique << ", " << s.leafWork << ", " << s.leafUnique << ", "
              << s.terminated << ", " << s.nuMax << ", " << s.nuMin << ", " << s.luMax << ", " << s.luMin << std::endl;
        }
    }
    {
        std::ofstream f = open(files[3]);
        f << "loadedPrimaryNodes, loadedPrimaryLeafs, loadedPrimaryNodesMax, loadedPrimaryLeafsMax, "
             "loadedPrimaryNodesMin, loadedPrimaryLeafsMin, loadedSecondaryNodes, loadedSecondaryLeafs, "
             "loadedSecondaryNodesMax, loadedSecondaryLeafsMax, loadedSecondaryNodesMin, loadedSecondaryLeafsMin, "
             "loadedWidePrimaryNodes, loadedWidePrimaryLeafs, loadedWideSecondaryNodes, loadedWideSecondaryLeafs"
          << std::endl;
        for (size_t i = 0; i < tiles; ++i) {
            float sums[4] = {0, 0, 0, 0};
            int mx[4] = {0, 0, 0, 0}, mn[4] = {1000, 1000, 1000, 1000};
            const std::vector<uint32_t>* v[4] = {&pv.nodes, &pv.leaves, &sv.nodes, &sv.leaves};
            for (unsigned j = 0; j < wgsq; ++j)
                for (int k = 0; k < 4; ++k) {
                    const int x = (int)(*v[k])[i * wgsq + j];
                    sums[k] += x;
                    mx[k] = std::max(x, mx[k]);
                    mn[k] = std::min(x, mn[k]);
                }
            for (int k = 0; k < 4; ++k)
                if (mn[k] == 1000) mn[k] = 0;
            int wide[4] = {0, 0, 0, 0};
            for (const auto& r : primary[i]) {
                wide[0] += (int)r.node_unique;
                wide[1] += (int)r.leaf_unique;
            }
            for (const auto& r : secondary[i]) {
                wide[2] += (int)r.node_unique;
                wide[3] += (int)r.leaf_unique;
            }
            f << sums[0] << ", " << sums[1] << ", " << mx[0] << ", " << mx[1] << ", " << mn[0] << ", " << mn[1] << ", "
              << sums[2] << ", " << sums[3] << ", " << mx[2] << ", " << mx[3] << ", " << mn[2] << ", " << mn[3] << ", "
              << (float)wide[0] << ", " << (float)wide[1] << ", " << (float)wide[2] << ", " << (float)wide[3]
              << std::endl;
        }
    }
    return files;
}

// ---------------------------------------------------------------------------

uint8_t ambient_gray(bool hit, unsigned occluded, unsigned samples) {
    if (!hit || samples == 0) return 0;
    const float factor = 1.0f - ((float)occluded / (float)samples);
    return (uint8_t)(factor * 255.0f);
}

std::vector<uint8_t> tile_order_to_rows(const std::vector<uint8_t>& tiles, unsigned width, unsigned height,
                                        unsigned wg, unsigned channels) {
    if (wg == 0 || width % wg != 0) throw Error("tile_order_to_rows: width must be a multiple of the workgroup size");
    if (tiles.size() != (size_t)width * height * channels) throw Error("tile_order_to_rows: wrong pixel count");
    const size_t tilesX = width / wg, fullRows = height / wg, rowsLeft = height - fullRows * wg;
    std::vector<uint8_t> out(tiles.size());
    for (size_t h = 0; h < height; ++h)
        for (size_t w = 0; w < width; ++w) {
            const size_t lane = (h % wg) * wg + w % wg;
            size_t k;
            if (h < fullRows * wg) k = ((h / wg) * tilesX + w / wg) * wg * wg + lane;
            else k = fullRows * tilesX * wg * wg + (w / wg) * rowsLeft * wg + lane;
            for (unsigned c = 0; c < channels; ++c) out[(h * width + w) * channels + c] = tiles[k * channels + c];
        }
    return out;
}

namespace {

void put_chunk(std::string& png, const char* tag, const std::string& data) {
    auto be32 = [&png](uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) png.push_back((char)((v >> s) & 0xFF));
    };
    be32((uint32_t)data.size());
    const std::string body = std::string(tag, 4) + data;
    png += body;
    be32((uint32_t)crc32(0L, reinterpret_cast<const Bytef*>(body.data()), (uInt)body.size()));
}

}  // namespace

void write_png(const std::string& path, const std::vector<uint8_t>& rgba, unsigned width, unsigned height) {
    if (rgba.size() != (size_t)width * height * 4) throw Error("write_png: expected width * height * 4 bytes");
    std::string raw;
    raw.reserve((size_t)height * (1 + 4 * (size_t)width));
    for (unsigned r = 0; r < height; ++r) {
        raw.push_back('\0');  // filter type 0
        raw.append(reinterpret_cast<const char*>(rgba.data()) + (size_t)r * width * 4, (size_t)width * 4);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::string z(zlen, '\0');
    if (compress2(reinterpret_cast<Bytef*>(&z[0]), &zlen, reinterpret_cast<const Bytef*>(raw.data()),
                  (uLong)raw.size(), 6) != Z_OK)
        throw Error("write_png: zlib compression failed");
    z.resize(zlen);
    std::string ihdr;
    for (uint32_t v : {width, height})
        for (int s = 24; s >= 0; s -= 8) ihdr.push_back((char)((v >> s) & 0xFF));
    ihdr += std::string("\x08\x06\x00\x00\x00", 5);  // 8 bit, RGBA, deflate, no filter, no interlace
    std::string png("\x89PNG\r\n\x1a\n", 8);
    put_chunk(png, "IHDR", ihdr);
    put_chunk(png, "IDAT", z);
    put_chunk(png, "IEND", "");
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error("write_png: cannot open " + path);
    f.write(png.data(), (std::streamsize)png.size());
}

// ---------------------------------------------------------------------------

CameraFast::CameraFast(std::string path, std::string name, std::string problem, std::string problemPrefix,
                       int workGroupSize, bool saveDistance, bool wideRender, std::vector<Vec3> positions,
                       std::vector<Vec3> lookCenters, unsigned width, unsigned height, Vec3 upward, float focalLength)
    : path_(std::move(path)), name_(std::move(name)), problem_(std::move(problem)), prefix_(std::move(problemPrefix)),
      wg_(workGroupSize), saveDistance_(saveDistance), wideRender_(wideRender), positions_(std::move(positions)),
      targets_(std::move(lookCenters)), width_(width), height_(height), up_(upward), focal_(focalLength) {
    if (positions_.size() != targets_.size()) throw Error("CameraFast: one look center per camera position");
    if (wg_ <= 0 || width_ % (unsigned)wg_ != 0) throw Error("CameraFast: width must be a multiple of workGroupSize");
}

std::string CameraFast::perfFile(bool wideAlternative) const {
    std::string wgName = "/WorkGroupSize_" + std::to_string(wg_) + "_Normal";
    if (wideRender_) wgName = "/WorkGroupSize_" + std::to_string(wg_) + (wideAlternative ? "_Wide" : "_WideOld");
    return path_ + wgName + "/" + name_ + problem_ + prefix_ + "_Perf.txt";
}

std::string CameraFast::imageFile(int cameraId) const {
    return path_ + "/" + name_ + "_c" + std::to_string(cameraId) + "_Perf.png";
}

std::tuple<float, float, float> CameraFast::renderImage(bool saveImage, FastNodeManager& m,
                                                        unsigned ambientSampleCount, float ambientDistance,
                                                        int cameraId, std::vector<uint8_t>* image) {
    rt_camera cam{};
    for (int a = 0; a < 3; ++a) {
        cam.pos[a] = positions_[cameraId][a];
        cam.target[a] = targets_[cameraId][a];
        cam.up[a] = up_[a];
    }
    cam.focal = focal_;
    const size_t n = (size_t)width_ * height_;
    std::vector<uint8_t> gray(n);
    float ms[2] = {0, 0};
    // the ray-stream traversal has no save-distance form
    const uint32_t flags = wideRender_ ? RT_STREAM : (saveDistance_ ? RT_SAVE_DIST : RT_CLOSEST);
    check(rt_render_ambient(m.scene(), &cam, width_, height_, (uint32_t)wg_, flags, ambientSampleCount,
                            ambientDistance, gray.data(), ms, nullptr),
          "rt_render_ambient");
    if (saveImage) {
        const std::vector<uint8_t> rows = tile_order_to_rows(gray, width_, height_, (unsigned)wg_);
        std::vector<uint8_t> rgba(n * 4);
        for (size_t i = 0; i < n; ++i) {
            rgba[4 * i] = rgba[4 * i + 1] = rgba[4 * i + 2] = rows[i];
            rgba[4 * i + 3] = 255;
        }
        write_png(imageFile(cameraId), rgba, width_, height_);
    }
    if (image) *image = std::move(gray);
    return std::make_tuple(ms[1] / 1000.0f, ms[0] / 1000.0f, 0.0f);
}

std::pair<float, float> CameraFast::renderImages(bool saveImage, FastNodeManager& m, unsigned ambientSampleCount,
                                                 float ambientDistance, bool mute, bool wideAlternative) {
    const int cameras = (int)positions_.size();
    if (cameras == 0) throw Error("CameraFast: no camera");
    std::error_code ec;
    std::filesystem::create_directories(path_, ec);
    if (saveImage)
        for (int c = 0; c < cameras; ++c) renderImage(true, m, ambientSampleCount, ambientDistance, c);
    constexpr int kSamples = 5;  // odd, for the median
    float totalTime = 0, timeSumEachRay = 0, timeSumTriangle = 0;
    for (int c = 0; c < cameras; ++c) {
        renderImage(false, m, ambientSampleCount, ambientDistance, c);  // warm-up
        std::vector<std::tuple<float, float, float>> res;
        for (int i = 0; i < kSamples; ++i) res.push_back(renderImage(false, m, ambientSampleCount, ambientDistance, c));
        std::sort(res.begin(), res.end());
        totalTime += std::get<0>(res[kSamples / 2]);
        timeSumEachRay += std::get<1>(res[kSamples / 2]);
        timeSumTriangle += std::get<2>(res[kSamples / 2]);
    }
    totalTime /= (float)cameras;
    timeSumEachRay /= (float)cameras;
    timeSumTriangle /= (float)cameras;
    if (!mute) {
        std::cout << "Raytracing took " << totalTime << " seconds." << std::endl;
        std::cout << "ray sum took " << timeSumEachRay << " seconds." << std::endl;
    }
    const std::string file = perfFile(wideAlternative);
    std::filesystem::create_directories(std::filesystem::path(file).parent_path(), ec);
    write_perf(file, name_, m.branchingFactor, m.leafSize, m.gangSize, m.nodeMemory, m.leafMemory, totalTime,
               timeSumEachRay, timeSumTriangle, m.averageBvhDepth, m.triangleCount);
    return {totalTime, timeSumEachRay};
}

void write_perf(const std::string& path, const std::string& name, unsigned branchingFactor, unsigned leafSize,
                unsigned gangSize, unsigned nodeMemory, unsigned leafMemory, float totalTime, float raySum,
                float triangleSum, float averageBvhDepth, uint64_t triangleCount) {
    std::ofstream f(path);
    if (!f.is_open()) throw Error("write_perf: cannot open " + path);
    f << "scenario " << name << " with branching factor of " << branchingFactor << " and leafsize of " << leafSize
      << std::endl;
    f << "with a gang size of " << gangSize << ", a node meory of " << nodeMemory << ", and a leaf memory of "
      << leafMemory << std::endl;
    f << "Raytracer total time: " << totalTime << std::endl;
    f << "Time for all rays (SUM): " << raySum << std::endl;
    f << "Time for triangle intersections (SUM): " << triangleSum << std::endl;
    f << "Time all rays(sum) - triangle(sum): " << raySum - triangleSum << std::endl;
    f << "Average BVH depth: " << averageBvhDepth << std::endl;
    f << "Triangle Count: " << triangleCount << std::endl;
}

// ---------------------------------------------------------------------------

CameraData::CameraData(std::string path, std::string name, std::string problem, int workGroupSize, bool wideRender,
                       std::vector<Vec3> positions, std::vector<Vec3> lookCenters, unsigned width, unsigned height,
                       Vec3 upward, float focalLength)
    : path_(std::move(path)), name_(std::move(name)), problem_(std::move(problem)), wg_(workGroupSize),
      wideRender_(wideRender), positions_(std::move(positions)), targets_(std::move(lookCenters)), width_(width),
      height_(height), up_(upward), focal_(focalLength) {
    if (positions_.size() != targets_.size()) throw Error("CameraData: one look center per camera position");
    if (wg_ <= 0 || width_ % (unsigned)wg_ != 0) throw Error("CameraData: width must be a multiple of workGroupSize");
}

std::string CameraData::renderImages(FastNodeManager& m, unsigned ambientSampleCount, float ambientDistance,
                                     bool doWorkGroupAnalysis, bool wideAlternative) {
    if (positions_.empty()) throw Error("CameraData: no camera");
    std::error_code ec;
    std::filesystem::create_directories(path_, ec);
    rt_trace_hist hp{}, hs{};
    const size_t n = (size_t)width_ * height_;
    const unsigned wgsq = (unsigned)(wg_ * wg_);
    const size_t tiles = (size_t)(width_ / wg_) * (height_ / wg_);
    for (size_t c = 0; c < positions_.size(); ++c) {
        rt_camera cam{};
        for (int a = 0; a < 3; ++a) {
            cam.pos[a] = positions_[c][a];
            cam.target[a] = targets_[c][a];
            cam.up[a] = up_[a];
        }
        cam.focal = focal_;
        std::vector<rt_ray> rays(n);
        check(rt_camera_rays(&cam, width_, height_, (uint32_t)wg_, rays.data()), "rt_camera_rays");
        std::vector<rt_hit> hits;
        m.histograms(rays, false, hp, &hits);
        std::vector<float> nrm, pos;
        m.getSurfaceNormalPosition(rays, hits, nrm, pos);
        std::vector<rt_ray> amb(n);
        std::vector<uint32_t> src(n);
        for (unsigned a = 0; a < ambientSampleCount; ++a) {
            uint32_t k = 0;
            check(rt_spawn_ambient(nrm.data(), pos.data(), hits.data(), (uint32_t)n, a, ambientDistance, amb.data(),
                                   src.data(), &k),
                  "rt_spawn_ambient");
            m.histograms(std::vector<rt_ray>(amb.begin(), amb.begin() + k), true, hs);
        }
        if (wideRender_ && doWorkGroupAnalysis) {
            if (ambientSampleCount == 0) throw Error("CameraData: the workgroup analysis needs ambient samples");
            const std::vector<rt_ray> full(rays.begin(), rays.begin() + tiles * wgsq);
            TileVisits pv, sv;
            const auto primary = m.workgroupSteps(full, wgsq, false, &pv);
            std::vector<std::vector<rt_step_stat>> secondary(tiles);
            sv.nodes.assign(full.size(), 0);
            sv.leaves.assign(full.size(), 0);
            rt_ray nanRay;
            std::fill_n(reinterpret_cast<float*>(&nanRay), 10, std::nanf(""));
            for (unsigned a = 0; a < ambientSampleCount; ++a) {
                // a pixel without a hit gets a NaN ray (cameraData.h:495-499)
                uint32_t k = 0;
                check(rt_spawn_ambient(nrm.data(), pos.data(), hits.data(), (uint32_t)full.size(), a, ambientDistance,
                                       amb.data(), src.data(), &k),
                      "rt_spawn_ambient");
                std::vector<rt_ray> dense(full.size(), nanRay);
                for (uint32_t i = 0; i < k; ++i) dense[src[i]] = amb[i];
                TileVisits v;
                const auto recs = m.workgroupSteps(dense, wgsq, true, &v);
                for (size_t t = 0; t < tiles; ++t) secondary[t].insert(secondary[t].end(), recs[t].begin(), recs[t].end());
                for (size_t r = 0; r < full.size(); ++r) {
                    sv.nodes[r] += v.nodes[r];
                    sv.leaves[r] += v.leaves[r];
                }
            }
            write_workgroup_analysis(path_, name_, problem_, (int)c, width_, height_, (unsigned)wg_, primary, secondary,
                                     pv, sv, wideAlternative);
        }
    }
    const std::string info = path_ + "/" + name_ + problem_ + "_Info.txt";
    write_info(info, name_, m.branchingFactor, m.leafSize, m.desc().max_depth + 1, hp, hs);
    return info;
}

// ---------------------------------------------------------------------------

std::vector<float> scenario_triangles(int scenario, std::string& name, Vec3& pos, Vec3& target) {
    float* tris = nullptr;
    uint32_t n = 0;
    if (scenario == 0) {
        check(rt_scene_generate(RT_SCENE_SPHERES, 0, 50, 10, &tris, &n), "rt_scene_generate");
        name = "spheres";
        pos = {0.0f, 4.0f, 32.0f};
        target = {0.0f, 0.0f, 0.0f};
    } else if (scenario == 1) {
        check(rt_scene_generate(RT_SCENE_INTERIOR, 2, 1000000, 0, &tris, &n), "rt_scene_generate");
        name = "interior";
        pos = {0.0f, 5.0f, -14.0f};
        target = {0.0f, 3.0f, 0.0f};
    } else {
        throw Error("unknown scenario " + std::to_string(scenario) + " (0 = spheres, 1 = interior)");
    }
    std::vector<float> out(tris, tris + (size_t)n * 9);
    rt_free(tris);
    return out;
}

float ambient_distance(const std::vector<float>& tris) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i + 2 < tris.size(); i += 3)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], tris[i + a]);
            hi[a] = std::max(hi[a], tris[i + a]);
        }
    const float volume = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    return std::cbrt(volume) / 10.f;
}

RayTracer::RayTracer(RenderConfig cfg, std::string path, unsigned gangSize)
    : cfg_(std::move(cfg)), path_(std::move(path)), gang_(gangSize) {
    sanity_clamp(cfg_);
}

float RayTracer::renderImage(unsigned branchingFactor, unsigned leafSize, unsigned gangSize,
                             const std::vector<float>& tris, const std::vector<Vec3>& cameraPositions,
                             const std::vector<Vec3>& cameraTargets, const std::string& name, float sahFactor) {
    const std::string problem = "_b" + std::to_string(branchingFactor) + "_l" + std::to_string(leafSize);
    FastNodeManager m(tris, branchingFactor, leafSize, gangSize, sahFactor, cfg_.sortEachSplit, cfg_.leafSplitOption);
    const float ambientDistance = ambient_distance(tris);
    const unsigned samples = (unsigned)std::max(0, cfg_.ambientSampleCount);
    float total = 0;
    if (cfg_.doPerformanceTest) {
        // leaf memory (outer) and node memory (inner) from the minimum up to
        // 3 gangs more with the memory tests (rayTracer.cpp:41-52,630-640);
        // this engine's records hold at most 16 slots
        const unsigned minLm = m.leafMemory, minM = m.nodeMemory;
        const unsigned maxLm = cfg_.doLeafMemoryTest ? std::min(16u, minLm + 3 * gangSize) : minLm;
        const unsigned maxM = cfg_.doNodeMemoryTest ? std::min(16u, minM + 3 * gangSize) : minM;
        for (unsigned lm = minLm; lm <= maxLm; lm += gangSize)
            for (unsigned nm = minM; nm <= maxM; nm += gangSize) {
                const bool base = lm == minLm && nm == minM;
                std::unique_ptr<FastNodeManager> other;
                if (!base)
                    other.reset(new FastNodeManager(tris, branchingFactor, leafSize, gangSize, sahFactor,
                                                    cfg_.sortEachSplit, cfg_.leafSplitOption, 0, nm, lm));
                FastNodeManager& mm = base ? m : *other;
                const std::string prefix = "_mb" + std::to_string(mm.nodeMemory) + "_ml" + std::to_string(mm.leafMemory);
                CameraFast cam(path_, name, problem, prefix, cfg_.workGroupSize, cfg_.saveDistance, cfg_.wideRender,
                               cameraPositions, cameraTargets, (unsigned)cfg_.xRes, (unsigned)cfg_.yRes);
                const float t = cam.renderImages(cfg_.saveImage && base, mm, samples, ambientDistance, true,
                                                 cfg_.wideAlternative)
                                    .first;
                if (base) total = t;
            }
    }
    if (cfg_.renderAnalysisImage) {
        CameraData cam(path_, name, problem, cfg_.workGroupSize, cfg_.wideRender, cameraPositions, cameraTargets,
                       (unsigned)cfg_.xRes, (unsigned)cfg_.yRes);
        cam.renderImages(m, samples, ambientDistance, cfg_.doWorkGroupAnalysis, cfg_.wideAlternative);
    }
    return total;
}

std::vector<std::tuple<int, int, int, float>> RayTracer::run() {
    std::vector<std::tuple<int, int, int, float>> out;
    if (!cfg_.doPerformanceTest && !cfg_.renderAnalysisImage) return out;
    std::vector<int> scenarios = cfg_.scenarios;
    if (scenarios.empty()) scenarios.push_back(0);
    for (int sc : scenarios) {
        std::string name;
        Vec3 pos, tgt;
        const std::vector<float> tris = scenario_triangles(sc, name, pos, tgt);
        for (const auto& lb : sweep(cfg_)) {
            const int leaf = lb.first, branch = lb.second;
            const float sah = load_sah_factor(leaf, branch, (int)gang_);
            const float t = renderImage((unsigned)branch, (unsigned)leaf, gang_, tris, {pos}, {tgt}, name, sah);
            out.emplace_back(sc, branch, leaf, t);
        }
    }
    return out;
}

}  // namespace rt
