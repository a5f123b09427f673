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
#include <filesystem>
#include <fstream>
#include <iostream>

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
                                 float sahFactor, bool sortEachSplit, int leafSplitOption, unsigned threads) {
    if (tris.size() % 9 != 0) throw Error("FastNodeManager: triangle array is not a multiple of 9 floats");
    rt_build_params p{branching, leaf, gang, sortEachSplit ? 1u : 0u, leafSplitOption, sahFactor, threads};
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
    const auto t0 = std::chrono::steady_clock::now();
    rt_camera cam{};
    for (int a = 0; a < 3; ++a) {
        cam.pos[a] = positions_[cameraId][a];
        cam.target[a] = targets_[cameraId][a];
        cam.up[a] = up_[a];
    }
    cam.focal = focal_;
    const size_t n = (size_t)width_ * height_;
    std::vector<rt_ray> rays(n);
    check(rt_camera_rays(&cam, width_, height_, (uint32_t)wg_, rays.data()), "rt_camera_rays");
    std::vector<rt_hit> hits;
    // the ray-stream traversal has no save-distance form
    m.intersect(rays, hits, nullptr, saveDistance_ && !wideRender_, wideRender_);
    const float tPrimary = (float)seconds_since(t0);
    std::vector<uint16_t> occl(n, 0);
    if (ambientSampleCount) {
        std::vector<float> nrm, pos;
        m.getSurfaceNormalPosition(rays, hits, nrm, pos);
        std::vector<rt_ray> amb(n);
        std::vector<uint32_t> src(n);
        std::vector<uint8_t> occ;
        for (unsigned a = 0; a < ambientSampleCount; ++a) {
            uint32_t k = 0;
            check(rt_spawn_ambient(nrm.data(), pos.data(), hits.data(), (uint32_t)n, a, ambientDistance, amb.data(),
                                   src.data(), &k),
                  "rt_spawn_ambient");
            if (k == 0) break;
            std::vector<rt_ray> batch(amb.begin(), amb.begin() + k);
            m.intersectSecondary(batch, occ, nullptr, wideRender_);
            for (uint32_t i = 0; i < k; ++i) occl[src[i]] += occ[i];
        }
    }
    const float total = (float)seconds_since(t0);
    if (image || saveImage) {
        std::vector<uint8_t> gray(n);
        for (size_t i = 0; i < n; ++i) gray[i] = ambient_gray(hits[i].slot >= 0, occl[i], ambientSampleCount);
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
    }
    return std::make_tuple(total, tPrimary, 0.0f);
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
    std::ofstream f(file);
    if (!f.is_open()) throw Error("CameraFast: cannot open " + file);
    f << "scenario " << name_ << " with branching factor of " << m.branchingFactor << " and leafsize of " << m.leafSize
      << std::endl;
    f << "with a gang size of " << m.gangSize << ", a node meory of " << m.nodeMemory << ", and a leaf memory of "
      << m.leafMemory << std::endl;
    f << "Raytracer total time: " << totalTime << std::endl;
    f << "Time for all rays (SUM): " << timeSumEachRay << std::endl;
    f << "Time for triangle intersections (SUM): " << timeSumTriangle << std::endl;
    f << "Time all rays(sum) - triangle(sum): " << timeSumEachRay - timeSumTriangle << std::endl;
    f << "Average BVH depth: " << m.averageBvhDepth << std::endl;
    f << "Triangle Count: " << m.triangleCount << std::endl;
    return {totalTime, timeSumEachRay};
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
    const std::string prefix = "_mb" + std::to_string(m.nodeMemory) + "_ml" + std::to_string(m.leafMemory);
    CameraFast cam(path_, name, problem, prefix, cfg_.workGroupSize, cfg_.saveDistance, cfg_.wideRender,
                   cameraPositions, cameraTargets, (unsigned)cfg_.xRes, (unsigned)cfg_.yRes);
    return cam.renderImages(cfg_.saveImage, m, (unsigned)std::max(0, cfg_.ambientSampleCount), ambientDistance, true,
                            cfg_.wideAlternative)
        .first;
}

std::vector<std::tuple<int, int, int, float>> RayTracer::run() {
    std::vector<std::tuple<int, int, int, float>> out;
    if (!cfg_.doPerformanceTest) return out;
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
