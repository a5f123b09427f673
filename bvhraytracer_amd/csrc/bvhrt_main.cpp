// bvhrt: the reference's program entry (src/main.cpp, RayTracer::readConfig
// + run) over the GPU engine.
//
//   bvhrt [config.txt] [output dir]     performance test of config.txt
//   bvhrt --print-config [config.txt]   the parsed settings, one key=value a line
#include <cstring>
#include <iostream>

#include "../../include/rt_renderer.hpp"

namespace {

void print_config(const rt::RenderConfig& c) {
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
    for (const auto& kv : ints) std::cout << kv.first << "=" << kv.second << "\n";
    for (const auto& kv : bools) std::cout << kv.first << "=" << (kv.second ? "true" : "false") << "\n";
    auto list = [](const char* k, const std::vector<int>& v) {
        std::cout << k << "=";
        for (size_t i = 0; i < v.size(); ++i) std::cout << (i ? "," : "") << v[i];
        std::cout << "\n";
    };
    list("scenarios", c.scenarios);
    list("cacheSizes", c.cacheSizes);
    for (const auto& w : c.warnings) std::cout << "warning=" << w << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::strcmp(argv[1], "--print-config") == 0) {
            rt::RenderConfig c;
            const std::string path = argc >= 3 ? argv[2] : "config.txt";
            if (!rt::read_config(path, c)) {
                std::cerr << "cannot open " << path << std::endl;
                return 2;
            }
            print_config(c);
            return 0;
        }
        const std::string path = argc >= 2 ? argv[1] : "config.txt";
        const std::string out = argc >= 3 ? argv[2] : "perf";
        rt::RenderConfig c;
        if (!rt::read_config(path, c)) {
            std::cerr << "cannot open " << path << std::endl;
            return 2;
        }
        for (const auto& w : c.warnings) std::cerr << w << std::endl;
        rt::RayTracer tracer(c, out);
        for (const auto& r : tracer.run())
            std::cout << "scenario " << std::get<0>(r) << " N" << std::get<1>(r) << " L" << std::get<2>(r)
                      << " median render " << std::get<3>(r) << " s" << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "bvhrt: " << e.what() << std::endl;
        return 1;
    }
}
