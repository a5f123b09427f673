// rt_renderer.hpp — C++ host layer of the MI355X ray-query engine.
//
// The reference keeps its renderer in C++: RayTracer reads config.txt, sweeps
// the branching factor and leaf size, builds the BVH and hands it to
// CameraFast, which renders every camera (1 warm-up + 5 timed renders, the
// median kept) through FastNodeManager and writes _Perf.txt / _Perf.png
// (src/rayTracer.cpp:97-122,568-745,756-925, src/cameraFast.h:49-582,829-852).
// This header is that layer with the traversal on the GPU: FastNodeManager
// builds with the reference's SAH builder (rt_build) and uploads the FastNode
// arrays (rt_upload_bvh); its intersect* calls are rt_trace over whole ray
// batches.  Everything here calls only the C ABI of rt.h.  Errors throw
// rt::Error with the library's message.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "rt.h"

namespace rt {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

using Vec3 = std::array<float, 3>;

// The config.txt settings of RayTracer (rayTracer.h:17-72), with the values a
// key takes when config.txt does not set it.
struct RenderConfig {
    int minLeafSize = 4, maxLeafSize = 4, leafStep = 4;
    int minBranch = 4, maxBranch = 4, branchStep = 4;
    int subdivisionStart = 0, subdivisionEnd = 0, subdivisionStep = 1;
    int renderType = 0, ambientSampleCount = 1, workGroupSize = 16, leafSplitOption = 1;
    int xRes = 1920, yRes = 1088;
    bool saveImage = false, saveDepthDetailedImage = false, bvhAnalysis = false, saveBvhImage = false;
    bool sortEachSplit = true, renderAnalysisImage = false, castShadows = false, doPerformanceTest = true;
    bool doNodeMemoryTest = false, doLeafMemoryTest = false, saveRayTimes = false, doWorkGroupAnalysis = false;
    bool saveDistance = false, wideRender = false, wideAlternative = true, renderAllOptions = false;
    bool doCacheAnalysis = false;
    std::vector<int> scenarios, cacheSizes;
    std::vector<std::string> warnings;  // what the reference prints to std::cerr
};

// RayTracer::readConfig (rayTracer.cpp:756-877, configUtil.h:6-58): every
// non-empty line not starting with '#' sets each key it contains (substring
// match) from the text after its first '=' (std::stoi); booleans look for
// "true" / "false" anywhere in the line.  Returns false if the file cannot be
// opened (cfg keeps its values).
bool read_config(const std::string& path, RenderConfig& cfg);

// The settings as "key=value" lines: the integers and booleans (true /
// false) by their config.txt key, scenarios and cacheSizes comma-separated,
// then one "warning=<text>" line per warning.
std::string config_text(const RenderConfig& cfg);

// The settings sanity checks of RayTracer::run (rayTracer.cpp:104-122).
void sanity_clamp(RenderConfig& cfg);

// (leafSize, branchingFactor) pairs of the sweep loop (rayTracer.cpp:568-576).
std::vector<std::pair<int, int>> sweep(const RenderConfig& cfg);

// RayTracer::loadSahFactor (rayTracer.cpp:880-925); table "" = the built-in
// sahFactorConfigSse.txt values.
float load_sah_factor(int leafSize, int branchingFactor, int gangSize, const std::string& table = "");

// Visit counts per ray in tile order (rt_trace_steps' counts, summed over the
// ambient samples for the secondary rays).
struct TileVisits {
    std::vector<uint32_t> nodes, leaves;
};

// The flattened BVH on the GPU: FastNodeManager<gang, nodeMemory, wg>'s
// traversals (fastNodeManager.h:835-1493) over ray batches.
class FastNodeManager {
   public:
    // Bvh(prims, N, L, sortEachSplit, leafSplitOption) + recursiveBvh(sahFactor)
    // + the FastNodeManager flattener (rt_build), uploaded to the current device.
    // nodeMemory / leafMemory: 0 = roundup(N, G) / roundup(L, G), or a larger
    // multiple of G (the node and leaf memory tests).
    FastNodeManager(const std::vector<float>& tris, unsigned branchingFactor, unsigned leafSize, unsigned gangSize,
                    float sahFactor, bool sortEachSplit = true, int leafSplitOption = 1, unsigned threads = 0,
                    unsigned nodeMemory = 0, unsigned leafMemory = 0);
    ~FastNodeManager();
    FastNodeManager(const FastNodeManager&) = delete;
    FastNodeManager& operator=(const FastNodeManager&) = delete;

    // intersect (:1328-1418) or, with saveDistance, intersectSaveDistance
    // (:1232-1326); wide = the ray-stream schedule (intersectWideAlternative,
    // :835-1049), same per-ray results.
    void intersect(const std::vector<rt_ray>& rays, std::vector<rt_hit>& hits, std::vector<rt_counts>* counts = nullptr,
                   bool saveDistance = false, bool wide = false);
    // intersectSecondary (:1420-1493): occluded[i] = 1 when ray i hits anything.
    void intersectSecondary(const std::vector<rt_ray>& rays, std::vector<uint8_t>& occluded,
                            std::vector<rt_counts>* counts = nullptr, bool wide = false);
    // getSurfaceNormalPosition (:1699-1726): 3 floats per ray each, NaN for a miss.
    void getSurfaceNormalPosition(const std::vector<rt_ray>& rays, const std::vector<rt_hit>& hits,
                                  std::vector<float>& normals, std::vector<float>& positions) const;
    // The instrumented traversal (rt_trace_histograms): adds this batch's
    // counters to *hist (closest hit, or any hit).
    void histograms(const std::vector<rt_ray>& rays, bool anyHit, rt_trace_hist& hist,
                    std::vector<rt_hit>* hits = nullptr);
    // Per-step workload of tiles of `tile` rays (rt_trace_steps), one record
    // vector per tile; visits receives each ray's node / leaf visits.
    std::vector<std::vector<rt_step_stat>> workgroupSteps(const std::vector<rt_ray>& rays, unsigned tile, bool anyHit,
                                                          TileVisits* visits = nullptr);

    unsigned branchingFactor, leafSize, gangSize, nodeMemory, leafMemory;
    float averageBvhDepth;
    uint32_t triangleCount;
    const rt_bvh_desc& desc() const { return desc_; }
    rt_scene_t scene() const { return scene_; }

   private:
    rt_host_bvh_t host_ = nullptr;
    rt_scene_t scene_ = nullptr;
    rt_bvh_desc desc_{};
};

// _Info.txt of the instrumented renderer (cameraData.h:241-329) from the
// rt_trace_histograms counters of the primary and the secondary (ambient)
// rays: per-ray averages, success ratios, waste factor, child fullness
// 0..N and leaf fullness 0..L of the primary rays, visits per depth
// 0..bvhDepth-1.  Counts must be nonzero (the reference divides by them).
void write_info(const std::string& path, const std::string& name, unsigned branchingFactor, unsigned leafSize,
                unsigned bvhDepth, const rt_trace_hist& primary, const rt_trace_hist& secondary);

// The workgroup analysis files of CameraData::workGroupAnalysisPerImage
// (cameraData.h:660-1109) under <path>/WorkGroupSize_<wg>_Version_<0|1>/:
// _PrimaryWorkGroupWiskerPlot.txt, _SecondaryWorkGroupWiskerPlot.txt,
// _WorkGroupData.txt, _WorkGroupUniqueWork.txt.  primary / secondary: the
// rt_trace_steps records of every full tile (secondary: every ambient
// sample's records back to back).  Returns the four paths.  Whisker rows
// that tie on (median, max) keep the tile order (std::sort leaves it open).
std::vector<std::string> write_workgroup_analysis(const std::string& path, const std::string& name,
                                                  const std::string& problem, int cameraId, unsigned width,
                                                  unsigned height, unsigned wg,
                                                  const std::vector<std::vector<rt_step_stat>>& primary,
                                                  const std::vector<std::vector<rt_step_stat>>& secondary,
                                                  const TileVisits& primaryVisits, const TileVisits& secondaryVisits,
                                                  bool wideAlternative = true);

// _Perf.txt of CameraFast::renderImages (cameraFast.h:147-161): the scenario
// and tree settings, the median times in seconds, the average BVH depth and
// the triangle count (floats streamed with the default 6 significant digits).
void write_perf(const std::string& path, const std::string& name, unsigned branchingFactor, unsigned leafSize,
                unsigned gangSize, unsigned nodeMemory, unsigned leafMemory, float totalTime, float raySum,
                float triangleSum, float averageBvhDepth, uint64_t triangleCount);

// Gray value of an ambient-occlusion pixel (cameraFast.h:404-439): a hit is
// (uint8)((1 - occluded / samples) * 255) in float32, a miss 0; with no
// samples every pixel is 0 (the reference converts NaN, 0 on x86).
uint8_t ambient_gray(bool hit, unsigned occluded, unsigned samples);

// Per-pixel values in CameraFast tile order to row-major (cameraFast.h:
// 833-849; a partial last tile row holds height % wg rows).  `channels`
// values per pixel.
std::vector<uint8_t> tile_order_to_rows(const std::vector<uint8_t>& tiles, unsigned width, unsigned height,
                                        unsigned wg, unsigned channels = 1);

// 8-bit RGBA PNG of height x width x 4 bytes (encodeTwoSteps, util.h:52-61).
void write_png(const std::string& path, const std::vector<uint8_t>& rgba, unsigned width, unsigned height);

// CameraFast (cameraFast.h:29-582): the performance renderer.
class CameraFast {
   public:
    CameraFast(std::string path, std::string name, std::string problem, std::string problemPrefix,
               int workGroupSize, bool saveDistance, bool wideRender, std::vector<Vec3> positions,
               std::vector<Vec3> lookCenters, unsigned width = 1920, unsigned height = 1088, Vec3 upward = {0, 1, 0},
               float focalLength = 0.866f);

    // renderImages (:50-161): with saveImage one untimed render per camera
    // writes <path>/<name>_c<id>_Perf.png; then per camera 1 warm-up and 5
    // timed renders, the median by total time kept, averaged over cameras and
    // written to <path>/WorkGroupSize_<wg>_<Normal|Wide|WideOld>/
    // <name><problem><problemPrefix>_Perf.txt.  Returns (total, ray sum).
    std::pair<float, float> renderImages(bool saveImage, FastNodeManager& m, unsigned ambientSampleCount,
                                         float ambientDistance, bool mute = true, bool wideAlternative = true);

    // renderImage (:170-582): primary rays (closest hit, or save distance),
    // hit surfaces, ambientSampleCount any-hit rays per hit, all on the GPU
    // (rt_render_ambient).  Returns (frame seconds, primary-ray seconds, 0) by
    // HIP events; image (optional) receives the gray value per pixel in tile
    // order; saveImage writes the PNG.
    std::tuple<float, float, float> renderImage(bool saveImage, FastNodeManager& m, unsigned ambientSampleCount,
                                                float ambientDistance, int cameraId,
                                                std::vector<uint8_t>* image = nullptr);

    std::string perfFile(bool wideAlternative) const;
    std::string imageFile(int cameraId) const;

   private:
    std::string path_, name_, problem_, prefix_;
    int wg_;
    bool saveDistance_, wideRender_;
    std::vector<Vec3> positions_, targets_;
    unsigned width_, height_;
    Vec3 up_;
    float focal_;
};

// CameraData's analysis renderer (cameraData.h:84-560) on the GPU: per camera
// the primary rays (closest hit) and ambientSampleCount ambient rays per hit
// go through the instrumented traversal, and _Info.txt gets their counters;
// with a wide renderer and doWorkGroupAnalysis the per-step tile workload of
// the primary and ambient rays (a pixel without a hit gets a NaN ray that is
// not started) goes into the workgroup analysis files.
class CameraData {
   public:
    CameraData(std::string path, std::string name, std::string problem, int workGroupSize, bool wideRender,
               std::vector<Vec3> positions, std::vector<Vec3> lookCenters, unsigned width = 1920,
               unsigned height = 1088, Vec3 upward = {0, 1, 0}, float focalLength = 0.866f);
    // Returns the _Info.txt path.
    std::string renderImages(FastNodeManager& m, unsigned ambientSampleCount, float ambientDistance,
                             bool doWorkGroupAnalysis, bool wideAlternative = true);

   private:
    std::string path_, name_, problem_;
    int wg_;
    bool wideRender_;
    std::vector<Vec3> positions_, targets_;
    unsigned width_, height_;
    Vec3 up_;
    float focal_;
};

// RayTracer (rayTracer.h:15-85) with the performance test: for every
// scenario and every (L, N) of the sweep, build, upload and render.
// Scenarios are procedural (the reference's glTF models are not part of its
// repository): 0 = 50 icospheres + ground plane (100,002 triangles, camera
// (0,4,32) -> (0,0,0)); 1 = a 1M-triangle room interior (camera (0,5,-14) ->
// (0,3,0)).
class RayTracer {
   public:
    explicit RayTracer(RenderConfig cfg, std::string path = "perf", unsigned gangSize = 4);
    // Returns the median total render time per (scenario, N, L).
    std::vector<std::tuple<int, int, int, float>> run();
    // RayTracer::renderImage (rayTracer.cpp:588-745): the performance test
    // (doPerformanceTest) and the analysis render (renderAnalysisImage).
    float renderImage(unsigned branchingFactor, unsigned leafSize, unsigned gangSize, const std::vector<float>& tris,
                      const std::vector<Vec3>& cameraPositions, const std::vector<Vec3>& cameraTargets,
                      const std::string& name, float sahFactor);
    const RenderConfig& config() const { return cfg_; }

   private:
    RenderConfig cfg_;
    std::string path_;
    unsigned gang_;
};

// ambientDistance of RayTracer::renderImage (rayTracer.cpp:607): cbrt of the
// root box volume / 10, in float, the root box being the triangles' bounds.
float ambient_distance(const std::vector<float>& tris);

// Triangles of a procedural scenario (see RayTracer), 9 floats each.
std::vector<float> scenario_triangles(int scenario, std::string& name, Vec3& cameraPos, Vec3& cameraTarget);

}  // namespace rt
