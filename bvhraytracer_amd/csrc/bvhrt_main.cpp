// bvhrt: the reference's program entry (src/main.cpp, RayTracer::readConfig
// + run) over the GPU engine.
//
//   bvhrt [config.txt] [output dir]     performance test of config.txt
//   bvhrt --print-config [config.txt]   the parsed settings, one key=value a line
#include <cstring>
#include <iostream>

#include "../../include/rt_renderer.hpp"

namespace {

void print_config(const rt::RenderConfig& c) { std::cout << rt::config_text(c); }

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
