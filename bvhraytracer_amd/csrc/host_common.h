// Internal host-side declarations shared by the librt host sources.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rt.h"

namespace rt {

// PCG32 (O'Neill, PCG-XSH-RR 64/32).  Portable and fully specified, unlike the
// std::*_distribution family whose output is implementation defined.
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t stream = 0) {
        state = 0;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    // [0,1) with 24 bits: (u >> 8) * 2^-24
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    double uniform_d() { return (double)(next() >> 8) * (1.0 / 16777216.0); }
    double normal() {  // Box-Muller, first variate only
        double u1 = uniform_d(), u2 = uniform_d();
        if (u1 < 1e-12) u1 = 1e-12;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct V3 {
    float x, y, z;
};

thread_local extern std::string g_last_error;
int set_error(int code, const std::string& msg);

// Host BVH: the reference FastNode arrays plus the data the hit
// reconstruction needs.
struct HostBvh {
    uint32_t N = 0, L = 0, G = 0, M = 0, Lm = 0;
    std::vector<float> bounds;        // node_count * 6M
    std::vector<uint32_t> base;       // node_count
    std::vector<int8_t> order;        // node_count * 3M
    std::vector<uint16_t> childType;  // node_count
    std::vector<float> tri;           // trianglePoints
    std::vector<int32_t> triIds;      // leaves * Lm
    float averageDepth = 0;
    uint32_t triangleCount = 0;
    uint32_t maxDepth = 0;
    uint32_t nodeCount() const { return (uint32_t)base.size(); }
};

}  // namespace rt
