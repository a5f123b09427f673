// Per-ray wide-BVH traversal kernels for gfx950 (see rt_kernels.h).
//
// Execution model: persistent waves.  Every lane owns one ray at a time and
// walks its own depth-first stack; when a lane's ray finishes, the wave
// refills all idle lanes with one atomicAdd on the global ray counter
// (wave-aggregated, ballot + mbcnt).  The top `S` stack entries of every
// lane live in LDS (column `tid`, row `sp % S`, so a wave's 64 lanes always
// hit 64 distinct banks); deeper entries spill to a per-lane global region.
//
// All float arithmetic is FP32 with contraction disabled (-ffp-contract=off)
// and correctly rounded division, in the order of rayTracerTemplate.ispc;
// max/min are written as the sse4 maxps/minps selects the ISPC code lowers
// to, so NaN and signed-zero cases resolve as in the reference.
#include <type_traits>

#include "rt_kernels.h"

namespace rt {
namespace {

struct Ray {
    float tmax, ox, oy, oz, ix, iy, iz, dx, dy, dz;
};

__device__ __forceinline__ float smax(float a, float b) { return (a > b) ? a : b; }
__device__ __forceinline__ float smin(float a, float b) { return (a < b) ? a : b; }

// aabbIntersect body for one slot (rayTracerTemplate.ispc:190-230).
__device__ __forceinline__ bool slab(float bnx, float bny, float bnz, float bxx, float bxy, float bxz,
                                     const Ray& r, float& t0out) {
    float t0 = 0.0f, t1 = r.tmax;
    float tnx = (bnx - r.ox) * r.ix, tny = (bny - r.oy) * r.iy, tnz = (bnz - r.oz) * r.iz;
    float tfx = (bxx - r.ox) * r.ix, tfy = (bxy - r.oy) * r.iy, tfz = (bxz - r.oz) * r.iz;
    float tmp;
    if (tnx > tfx) { tmp = tnx; tnx = tfx; tfx = tmp; }
    t0 = smax(tnx, t0);
    t1 = smin(tfx, t1);
    if (tny > tfy) { tmp = tny; tny = tfy; tfy = tmp; }
    t0 = smax(tny, t0);
    t1 = smin(tfy, t1);
    if (tnz > tfz) { tmp = tnz; tnz = tfz; tfz = tmp; }
    t0 = smax(tnz, t0);
    t1 = smin(tfz, t1);
    t0out = t0;
    return t0 <= t1;
}

// slab() for operands that cannot produce a NaN: finite inverse direction,
// |origin| and |box| <= 2^100 (so b - o is finite and (b - o) * inv is finite
// or infinite, never NaN) and tmax not NaN.  Without NaNs the maxps/minps
// selects and the swap equal fmaxf/fminf in value; only the sign of a zero
// t0 may differ, and t0 is only ever compared.
__device__ __forceinline__ bool slab_fast(float bnx, float bny, float bnz, float bxx, float bxy, float bxz,
                                          const Ray& r, float& t0out) {
    const float tnx = (bnx - r.ox) * r.ix, tny = (bny - r.oy) * r.iy, tnz = (bnz - r.oz) * r.iz;
    const float tfx = (bxx - r.ox) * r.ix, tfy = (bxy - r.oy) * r.iy, tfz = (bxz - r.oz) * r.iz;
    const float t0 = fmaxf(fmaxf(fminf(tnx, tfx), fminf(tny, tfy)), fmaxf(fminf(tnz, tfz), 0.0f));
    const float t1 = fminf(fminf(fmaxf(tnx, tfx), fmaxf(tny, tfy)), fminf(fmaxf(tnz, tfz), r.tmax));
    t0out = t0;
    return t0 <= t1;
}

// Per-ray condition for slab_fast (the box bound is checked per scene) and
// for skipping a leaf's trailing all-zero triangles: with a finite direction
// such a triangle has e1 = e2 = 0, so s1 = 0 and divisor = 0, and it never
// passes the test (an infinite direction component would make the divisor
// NaN, which passes).
// Q: the quantized records' condition, which also bounds the origin and the
// inverse direction by 2^60 (their boxes are within 2^60, so slab_fast's
// condition holds too).
template <bool Q = false>
__device__ __forceinline__ bool ray_fast(const Ray& r) {
    constexpr float kMaxO = Q ? kQMaxCoord : 0x1p100f, kMaxF = 3.40282347e38f, kMaxI = Q ? kQMaxCoord : kMaxF;
    return fabsf(r.ox) <= kMaxO && fabsf(r.oy) <= kMaxO && fabsf(r.oz) <= kMaxO && fabsf(r.ix) <= kMaxI &&
           fabsf(r.iy) <= kMaxI && fabsf(r.iz) <= kMaxI && fabsf(r.dx) <= kMaxF && fabsf(r.dy) <= kMaxF &&
           fabsf(r.dz) <= kMaxF && r.tmax == r.tmax;
}

// 1/x correctly rounded, as IEEE division gives it.  For |x| in [2^-126,
// 2^126) the hardware reciprocal refined by one Newton-Raphson step,
// y + y * (1 - x*y) with fused multiply-adds, equals the correctly rounded
// quotient for every float (checked exhaustively on gfx950: all 4.2e9 values
// with |x| in [2^-126, 2^126] agree bit for bit with the division); zero,
// denormal, huge, infinite and NaN divisors take the full division.
__device__ __forceinline__ float recip(float x) {
    const uint32_t e = (__float_as_uint(x) >> 23) & 0xFFu;
    if (e - 1u < 252u) {
        const float y = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
    }
    return 1.0f / x;
}

// Result of triIntersect's gang rule from the mask of slots at the minimum
// t: a slot is its program instance's kept slot when no slot G, 2G, ... below
// it is in the mask (the instance met the minimum there first); the largest
// kept slot wins (rayTracerTemplate.ispc:80-101).
__device__ __forceinline__ int gang_slot(uint32_t hm, uint32_t G) {
    uint32_t below = 0;
    for (uint32_t sh = G; sh < 16; sh += G) below |= hm << sh;
    return 31 - (int)__builtin_clz(hm & ~below);
}

// Moeller-Trumbore body of triIntersect / triAnyHit (rayTracerTemplate.ispc:49-75).
// The edge vectors e1 = p1 - p0 and e2 = p2 - p0 come precomputed from the
// leaf block (same FP32 subtraction, done once at upload).
__device__ __forceinline__ bool mt(float p0x, float p0y, float p0z, float e1x, float e1y, float e1z,
                                   float e2x, float e2y, float e2z, const Ray& r, float& t,
                                   float& b1, float& b2) {
    const float s1x = r.dy * e2z - r.dz * e2y;
    const float s1y = r.dz * e2x - r.dx * e2z;
    const float s1z = r.dx * e2y - r.dy * e2x;
    const float divisor = s1x * e1x + s1y * e1y + s1z * e1z;
    bool hit = !(divisor == 0.0f);
    const float inv = recip(divisor);
    const float dx = r.ox - p0x, dy = r.oy - p0y, dz = r.oz - p0z;
    b1 = (dx * s1x + dy * s1y + dz * s1z) * inv;
    if (b1 < 0.0f || b1 > 1.0f) hit = false;
    const float s2x = dy * e1z - dz * e1y;
    const float s2y = dz * e1x - dx * e1z;
    const float s2z = dx * e1y - dy * e1x;
    b2 = (r.dx * s2x + r.dy * s2y + r.dz * s2z) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) hit = false;
    t = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
    if (t < 0.0f || t > r.tmax) hit = false;
    return hit;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ uint4 ld4u(const uint32_t* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ float comp(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

// Triangles of a leaf are read a pair of slots at a time: the pair's
// 80-byte chunk (the leaf block layout) as five 16-byte loads, the last
// carrying chunk 0's count of slots to test.
constexpr uint32_t LEAF_CHUNK = 2;
struct LeafChunk {
    float4 q[5];
    __device__ __forceinline__ void load(const float* T, uint32_t s0) {
        const float* c = T + kLeafHead + (s0 >> 1) * 20;
#pragma unroll
        for (int i = 0; i < 5; ++i) q[i] = ld4(c + 4 * i);
    }
    // value i (0..8) of slot s0 + j
    __device__ __forceinline__ float v(int i, int j) const { return comp(q[(2 * i + j) >> 2], (2 * i + j) & 3); }
    // chunk 0 only: slots to test for a finite direction (the layout's word 18)
    __device__ __forceinline__ uint32_t test_slots() const { return __float_as_uint(q[4].z); }
};

// maxAbsDimension (glmUtil.h:18-23)
__device__ __forceinline__ int max_abs_dim(float x, float y, float z) {
    x = fabsf(x);
    y = fabsf(y);
    z = fabsf(z);
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

// Header fields of an inner node visit (the layout's header block) for a ray
// whose major axis is `axis`: one 16-byte load when M <= 8.
struct NodeHeader {
    uint32_t innerBase, leafBase, leafMask, c, kcodes;
    uint64_t row;  // the axis' order row, 4-bit slots
};

__device__ __forceinline__ NodeHeader load_header(const uint32_t* rec, uint32_t M, int axis) {
    NodeHeader h;
    const uint32_t* hw = rec + 6 * M;
    if (M == 4) {
        const uint4 v = ld4u(hw);
        h.innerBase = v.x & kEntryIndexMask;
        h.leafMask = v.x >> 28;
        h.leafBase = v.y & kEntryIndexMask;
        h.c = (v.y >> 28) + 1u;
        h.kcodes = v.z & 0xFFu;
        h.row = ((((uint64_t)v.w << 32) | v.z) >> (8 + 16 * axis)) & 0xFFFFu;
    } else if (M == 8) {
        const uint4 v = ld4u(hw + 4 * axis);
        h.innerBase = v.x;
        h.leafBase = v.y;
        h.leafMask = v.z & 0xFFu;
        h.c = (v.z >> 8) & 15u;
        h.kcodes = v.z >> 16;
        h.row = v.w;
    } else {
        // M = 12, 16: two order words per axis after childBase (w4)
        const uint4 h0 = ld4u(hw), h1 = ld4u(hw + 4), h2 = ld4u(hw + 8);
        h.innerBase = h0.x;
        h.leafBase = h0.y;
        h.leafMask = h0.z & 0xFFFFu;
        h.c = (h0.z >> 16) & 31u;
        h.kcodes = h0.w;
        const uint32_t lo = axis == 0 ? h1.y : (axis == 1 ? h1.w : h2.y);
        const uint32_t hi = axis == 0 ? h1.z : (axis == 1 ? h2.x : h2.z);
        h.row = ((uint64_t)hi << 32) | lo;
    }
    return h;
}

// ---------------------------------------------------------------------------
// Traversal state of one ray and the step function shared by the per-ray
// kernel (state in registers) and the ray-stream kernel (state in LDS).

struct SceneCtx {
    const uint32_t* nodes;
    const float* tris;
    const int32_t* triIds;
    const uint32_t* nodeEntry;
    const uint8_t* leafPrims;  // real triangles per leaf block (visit statistics)
    uint32_t rootEntry, G, Lm, nodeCount, M, recw;
    bool boxesBounded;  // every tested box coordinate is within +-2^100
};

struct StackCtx {
    uint32_t* stE;   // LDS stack entries of this ray, row stride `stride`
    float* stD;      // LDS stack distances (save dist); tmax at the push
                     // (quantized records)
    float* aabbD;    // LDS aabbDistances (save dist)
    uint32_t* spE;   // global spill of this ray
    float* spD;
    uint32_t cap;    // stack capacity (LDS window + spill)
    uint32_t stride;
    // visit statistics only: depth of each stack entry (LDS window and
    // spill, laid out like stE/spE) and the block's LDS counters
    uint32_t* stDep;
    uint32_t* spDep;
    uint32_t* hist;
};

struct Trav {
    Ray ray;
    uint32_t cur;     // entry visited next (the child pushed last)
    float curDist;
    uint32_t sp;      // entries below `cur` on the stack
    uint32_t live;    // slots whose aabbDistances entry is not NaN
    uint32_t hitEntry;
    int hitSlot;
    float b1, b2;     // barycentrics of the hit (when the step tracks them)
    uint32_t cN, cL, cNS, cLS;
    bool hasCur, occl, err, anyMode, rev, fast;
    int axis;
    uint32_t curDepth;  // visit statistics: depth of `cur` (root = 0)
    // quantized records: tmax at the push of `cur` (used when its entry is
    // tentative)
    float curT;
};

template <bool Q = false>
__device__ __forceinline__ void trav_init(Trav& t, const rt_ray& in, uint32_t rootEntry, bool anyMode,
                                          bool boxesBounded) {
    t.ray.tmax = in.tmax;
    t.ray.ox = in.ox;
    t.ray.oy = in.oy;
    t.ray.oz = in.oz;
    t.ray.ix = in.ix;
    t.ray.iy = in.iy;
    t.ray.iz = in.iz;
    t.ray.dx = in.dx;
    t.ray.dy = in.dy;
    t.ray.dz = in.dz;
    t.axis = max_abs_dim(t.ray.dx, t.ray.dy, t.ray.dz);
    const float da = t.axis == 0 ? t.ray.dx : (t.axis == 1 ? t.ray.dy : t.ray.dz);
    t.rev = da <= 0.0f;
    t.fast = boxesBounded && ray_fast<Q>(t.ray);
    t.cur = rootEntry;
    t.curDist = 0.0f;
    t.curDepth = 0;
    t.hasCur = true;
    t.sp = 0;
    t.live = 0;
    t.hitEntry = 0;
    t.hitSlot = -1;
    t.b1 = t.b2 = 0.0f;
    t.occl = false;
    t.err = false;
    t.anyMode = anyMode;
    t.cN = t.cL = t.cNS = t.cLS = 0;
    t.curT = t.ray.tmax;
}

// Entry the next step visits (without taking it).
template <int S>
__device__ __forceinline__ uint32_t trav_peek(const Trav& t, const StackCtx& k) {
    return t.hasCur ? t.cur : k.stE[((t.sp - 1) & (S - 1)) * k.stride];
}

__device__ __forceinline__ bool trav_done(const Trav& t) { return !t.hasCur && t.sp == 0; }

// Visit statistics (rt_trace_histograms): LDS counter words of a block.
constexpr uint32_t kHistDepthBins = 64, kHistFullBins = 17;
enum HistWord : uint32_t {
    kHwNodeDepth = 0,
    kHwLeafDepth = kHwNodeDepth + kHistDepthBins,
    kHwChildFull = kHwLeafDepth + kHistDepthBins,
    kHwLeafFull = kHwChildFull + kHistFullBins,
    kHwAabbTests = kHwLeafFull + kHistFullBins,
    kHwAabbHits,
    kHwTriTests,
    kHwTriHits,
    kHwWords
};

// One traversal step: take the next entry and visit it (node or leaf).
// HIST adds the visit statistics (depth per stack entry, LDS counters).
// TRACK keeps the closest hit's barycentrics and slot as the leaf test finds
// them (the finish then writes them without re-running the triangle test).
template <int MT, int MODE, int S, bool HIST = false, bool TRACK = false>
__device__ __forceinline__ void trav_step(const SceneCtx& sc, const StackCtx& sk, Trav& t) {
    const uint32_t M = MT ? (uint32_t)MT : sc.M;
    const uint32_t recw = MT ? record_words(MT) : sc.recw;
    const uint32_t bd = sk.stride;
    uint32_t e;
    float edist = 0.0f;
    uint32_t edep = 0;
    if (t.hasCur) {
        e = t.cur;
        edist = t.curDist;
        if constexpr (HIST) edep = t.curDepth;
        t.hasCur = false;
    } else {
        --t.sp;
        const uint32_t slot = (t.sp & (S - 1)) * bd;
        e = sk.stE[slot];
        if (MODE == kSaveDist) edist = sk.stD[slot];
        if constexpr (HIST) edep = sk.stDep[slot];
        if (t.sp >= S) {
            sk.stE[slot] = sk.spE[t.sp - S];
            if (MODE == kSaveDist) sk.stD[slot] = sk.spD[t.sp - S];
            if constexpr (HIST) sk.stDep[slot] = sk.spDep[t.sp - S];
        }
    }
    if (MODE == kSaveDist && edist > t.ray.tmax) return;
    if (e & kEntryInvalid) {
        t.err = true;
        return;
    }
    const uint32_t k = ((e >> kEntryKShift) & 3u) * 4u + 4u;
    const uint32_t idx = e & kEntryIndexMask;
    if (e & kEntryLeaf) {
        const float* T = sc.tris + (size_t)idx * leaf_stride(sc.Lm);
        ++t.cL;
        t.cLS += k;
        uint32_t prims = 0, triHits = 0;
        if constexpr (HIST) prims = sc.leafPrims[idx];
        // slots to test: all k, or for a finite direction up to the last
        // triangle that is not all zeros; chunk 0 is always tested and
        // carries that count
        uint32_t kt = k;
        if (t.anyMode) {
            bool any = false;
            for (uint32_t s0 = 0; s0 == 0 || s0 < kt; s0 += LEAF_CHUNK) {
                LeafChunk a;
                a.load(T, s0);
                if (s0 == 0 && t.fast) kt = min(k, a.test_slots());
#pragma unroll
                for (int j = 0; j < LEAF_CHUNK; ++j) {
                    float tt, b1, b2;
                    if (mt(a.v(0, j), a.v(1, j), a.v(2, j), a.v(3, j), a.v(4, j), a.v(5, j), a.v(6, j), a.v(7, j),
                           a.v(8, j), t.ray, tt, b1, b2)) {
                        any = true;
                        if constexpr (HIST) triHits += (s0 + (uint32_t)j < prims) ? 1u : 0u;
                    }
                }
            }
            if (any) {
                t.occl = true;
                t.sp = 0;
                t.hasCur = false;
            }
        } else {
            // gang-G tie rule of triIntersect: each program instance
            // (slot % G) keeps its first strict minimum (tStore starts at
            // 1e6), the result is the largest kept slot among instances at
            // the minimum.  Kept here as the minimum t and the mask of slots
            // at it; gang_slot picks the result from the mask.  With TRACK
            // the result is followed slot by slot instead: slots come in
            // increasing order, so a slot at the minimum whose instance has
            // no earlier slot there is the largest kept slot so far.
            float bt = 1000000.0f;
            uint32_t hm = 0;
            int win = -1;
            float wb1 = 0.0f, wb2 = 0.0f;
            // slots of one program instance: every G-th bit (G = 4, 8, 16)
            const uint32_t inst = TRACK ? 0xFFFFu / ((1u << sc.G) - 1u) : 0u;
            for (uint32_t s0 = 0; s0 == 0 || s0 < kt; s0 += LEAF_CHUNK) {
                LeafChunk a;
                a.load(T, s0);
                if (s0 == 0 && t.fast) kt = min(k, a.test_slots());
#pragma unroll
                for (int j = 0; j < LEAF_CHUNK; ++j) {
                    float tt, b1, b2;
                    const bool h = mt(a.v(0, j), a.v(1, j), a.v(2, j), a.v(3, j), a.v(4, j), a.v(5, j), a.v(6, j),
                                      a.v(7, j), a.v(8, j), t.ray, tt, b1, b2);
                    const uint32_t s = s0 + (uint32_t)j;
                    if constexpr (HIST) triHits += (h && s < prims) ? 1u : 0u;
                    const bool less = h && tt < bt;
                    const bool same = h && tt == bt && hm != 0;
                    if constexpr (TRACK) {
                        const bool take = less || (same && (hm & (inst << (s & (sc.G - 1u)))) == 0);
                        win = take ? (int)s : win;
                        wb1 = take ? b1 : wb1;
                        wb2 = take ? b2 : wb2;
                    }
                    bt = less ? tt : bt;
                    hm = less ? (1u << s) : (same ? (hm | (1u << s)) : hm);
                }
            }
            if (hm) {
                t.ray.tmax = bt;
                t.hitEntry = e;
                if constexpr (TRACK) {
                    t.hitSlot = win;
                    t.b1 = wb1;
                    t.b2 = wb2;
                } else {
                    t.hitSlot = gang_slot(hm, sc.G);
                }
            }
        }
        if constexpr (HIST) {
            atomicAdd(&sk.hist[kHwLeafDepth + min(edep, kHistDepthBins - 1)], 1u);
            atomicAdd(&sk.hist[kHwLeafFull + min(prims, kHistFullBins - 1)], 1u);
            atomicAdd(&sk.hist[kHwTriTests], prims);
            atomicAdd(&sk.hist[kHwTriHits], triHits);
        }
        return;
    }
    const uint32_t* rec = sc.nodes + (size_t)idx * recw;
    ++t.cN;
    t.cNS += k;
    const NodeHeader hd = load_header(rec, M, t.axis);
    uint32_t H = 0;
    const float* bp = reinterpret_cast<const float*>(rec);
    auto boxes = [&](auto fastTag) {
        for (uint32_t s0 = 0; s0 < k; s0 += 4) {
            const float4 nx = ld4(bp + s0), ny = ld4(bp + M + s0), nz = ld4(bp + 2 * M + s0);
            const float4 xx = ld4(bp + 3 * M + s0), xy = ld4(bp + 4 * M + s0), xz = ld4(bp + 5 * M + s0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float t0;
                const float bnx = comp(nx, j), bny = comp(ny, j), bnz = comp(nz, j);
                const float bxx = comp(xx, j), bxy = comp(xy, j), bxz = comp(xz, j);
                bool hit;
                if constexpr (decltype(fastTag)::value) hit = slab_fast(bnx, bny, bnz, bxx, bxy, bxz, t.ray, t0);
                else hit = slab(bnx, bny, bnz, bxx, bxy, bxz, t.ray, t0);
                if (hit) {
                    H |= 1u << (s0 + (uint32_t)j);
                    if (MODE == kSaveDist) sk.aabbD[(s0 + (uint32_t)j) * bd] = t0;
                }
            }
        }
    };
    if (t.fast) boxes(std::true_type{});
    else boxes(std::false_type{});
    t.live |= H;
    if constexpr (HIST) {
        const uint32_t cc = hd.c;
        atomicAdd(&sk.hist[kHwNodeDepth + min(edep, kHistDepthBins - 1)], 1u);
        atomicAdd(&sk.hist[kHwChildFull + min(cc, kHistFullBins - 1)], 1u);
        atomicAdd(&sk.hist[kHwAabbTests], cc);
        atomicAdd(&sk.hist[kHwAabbHits], (uint32_t)__popc(H & ((1u << cc) - 1u)));
    }
    if (MODE != kSaveDist && H == 0) return;
    const uint32_t innerBase = hd.innerBase, leafBase = hd.leafBase;
    const uint32_t leafMask = hd.leafMask, c = hd.c;
    const uint32_t kcodes = hd.kcodes;
    const uint64_t row = hd.row;
    // stack entry of a real child slot s < c (no branches: both child kinds
    // are computed and selected)
    const uint32_t innerMask = ~leafMask;
    auto entry_real = [&](uint32_t s) -> uint32_t {
        const uint32_t below = (1u << s) - 1u;
        const uint32_t kc = ((kcodes >> (2 * s)) & 3u) << kEntryKShift;
        const bool isLeaf = (leafMask >> s) & 1u;
        const uint32_t idx = (isLeaf ? leafBase : innerBase) + __popc((isLeaf ? leafMask : innerMask) & below);
        return (isLeaf ? kEntryLeaf : 0u) | kc | idx;
    };
    // push(e, s): the previous pending child goes to the stack, child s
    // (entry e) becomes the pending one (visited next).  The stack capacity
    // is checked once per node, before the pushes.
    auto push = [&](uint32_t e, uint32_t s) {
        if (t.hasCur) {
            const uint32_t sl = (t.sp & (S - 1)) * bd;
            if (t.sp >= S) {
                sk.spE[t.sp - S] = sk.stE[sl];
                if (MODE == kSaveDist) sk.spD[t.sp - S] = sk.stD[sl];
            }
            sk.stE[sl] = t.cur;
            if (MODE == kSaveDist) sk.stD[sl] = t.curDist;
            if constexpr (HIST) {
                if (t.sp >= S) sk.spDep[t.sp - S] = sk.stDep[sl];
                sk.stDep[sl] = t.curDepth;
            }
            ++t.sp;
        }
        t.cur = e;
        if constexpr (HIST) t.curDepth = edep + 1u;
        if (MODE == kSaveDist) {
            t.curDist = sk.aabbD[s * bd];
            sk.aabbD[s * bd] = __builtin_nanf("");
        }
        t.hasCur = true;
        t.live &= ~(1u << s);
    };
    // entries pushed after the first land on the stack
    // the highest stack position a node's pushes write is sp + npush - 1
    // (the previous pending entry first); positions run up to cap - 1
    auto fits = [&](uint32_t npush) { return npush == 0 || t.sp + npush <= sk.cap; };
    // The push sequence is the order row filtered by the live mask.  Each
    // slot occurs once in the row, so the positions to push are known before
    // the first push: the loops below run once per push, not once per slot.
    if (t.anyMode) {
        // intersectSecondary: slot order 0..M-1 (fastNodeManager.h:1481-1488),
        // padded slots s >= c included: their FastNode id resolves like the
        // reference's, through the entry table
        const uint32_t m0 = t.live;
        if (!fits((uint32_t)__popc(m0))) {
            t.err = true;
        } else {
            for (uint32_t m = m0; m; m &= m - 1u) {
                const uint32_t s = (uint32_t)__builtin_ctz(m);
                uint32_t e;
                if (s >= c) {
                    const uint32_t id = rec[6 * M + child_base_word(M)] + s;
                    e = id < sc.nodeCount ? sc.nodeEntry[id] : kEntryInvalid;
                } else {
                    e = entry_real(s);
                }
                push(e, s);
            }
        }
    } else {
        uint32_t live = t.live;
        // backward over the zero-padded order row: the padding (slot 0)
        // comes first when childCount < M
        const bool pad0 = !t.rev && c < M && (live & 1u);
        if (pad0) live &= ~1u;
        uint32_t pm = 0;  // row positions q < c whose slot is live
        if constexpr (MT > 0) {
            // unrolled over the node memory; positions >= c masked off
            const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
#pragma unroll
            for (uint32_t q = 0; q < (uint32_t)MT; ++q) {
                const uint32_t slotq = __builtin_amdgcn_ubfe(q < 8 ? lo : hi, 4 * (q & 7u), 4);
                pm |= __builtin_amdgcn_ubfe(live, slotq, 1) << q;
            }
            pm &= (1u << c) - 1u;
        } else {
            for (uint32_t q = 0; q < c; ++q) pm |= ((live >> ((uint32_t)(row >> (4 * q)) & 15u)) & 1u) << q;
        }
        const uint32_t npush = (uint32_t)__popc(pm) + (pad0 ? 1u : 0u);
        const uint32_t rowLo = (uint32_t)row, rowHi = (uint32_t)(row >> 32);
        if (MODE != kSaveDist && !HIST && __ballot(t.sp + npush > (uint32_t)S) == 0) {
            // Every push of the wave lands in the LDS windows, so each entry
            // is written straight to its stack position: push order o goes to
            // sp + o, and the entry of order npush - 1 becomes the pending
            // one.  Backward lanes push the row positions in increasing
            // order, forward lanes (after the padding slot) in decreasing
            // order; the loop takes the lowest position first for both and
            // steps the stack position up or down.  Row positions q < c hold
            // exactly the real slots 0..c-1 (a permutation, checked at
            // upload), so the pushes clear the live bits of slots 0..c-1.
            const uint32_t leafBaseF = kEntryLeaf | leafBase;
            auto entry_fast = [&](uint32_t s) -> uint32_t {
                const bool isLeaf = __builtin_amdgcn_ubfe(leafMask, s, 1) != 0;
                const uint32_t m = isLeaf ? leafMask : innerMask;
                const uint32_t b = isLeaf ? leafBaseF : innerBase;
                const uint32_t kc = __builtin_amdgcn_ubfe(kcodes, 2 * s, 2);
                return (kc << kEntryKShift) | (b + (uint32_t)__popc(__builtin_amdgcn_ubfe(m, 0, s)));
            };
            if (npush) {
                const uint32_t last = t.sp + npush - 1u;
                uint32_t lastE = 0;
                if (pad0) {
                    lastE = entry_fast(0);
                    sk.stE[t.sp * bd] = lastE;
                }
                uint32_t pos = t.rev ? t.sp : last;
                const uint32_t step = t.rev ? 1u : ~0u;
                while (pm) {
                    const uint32_t q = (uint32_t)__builtin_ctz(pm);
                    pm &= pm - 1u;
                    const uint32_t s = (MT > 0 && MT <= 8) ? __builtin_amdgcn_ubfe(rowLo, 4 * q, 4)
                                                           : __builtin_amdgcn_ubfe(q < 8 ? rowLo : rowHi, 4 * (q & 7u), 4);
                    const uint32_t e = entry_fast(s);
                    sk.stE[pos * bd] = e;
                    lastE = pos == last ? e : lastE;
                    pos += step;
                }
                t.cur = lastE;
                t.hasCur = true;
                t.sp = last;
            }
            t.live &= ~((1u << c) - 1u);
        } else if (!fits(npush)) {
            t.err = true;
        } else {
            if (pad0) push(entry_real(0), 0);
            // one loop for both directions (forward takes the lowest position
            // first, backward the highest), so lanes of either kind share
            // the iterations; row positions q < c hold real slots only
            while (pm) {
                const uint32_t q = t.rev ? (uint32_t)__builtin_ctz(pm) : 31u - (uint32_t)__builtin_clz(pm);
                pm &= ~(1u << q);
                const uint32_t s = (MT > 0 && MT <= 8) ? __builtin_amdgcn_ubfe(rowLo, 4 * q, 4)
                                                       : __builtin_amdgcn_ubfe(q < 8 ? rowLo : rowHi, 4 * (q & 7u), 4);
                push(entry_real(s), s);
            }
        }
    }
    if (t.err) {
        t.sp = 0;
        t.hasCur = false;
    }
}

// ---------------------------------------------------------------------------
// Traversal step over the quantized node records (closest and any hit).
//
// A node's real children are tested on their quantized planes (rt_layout.h).
// The test is conservative: for a ray with |o|, |inv| <= 2^60 (t.fast) its
// near distance is at most, and its far distance at least, what the
// reference computes on the exact box, so every child the reference's
// aabbIntersect accepts passes.  Such a child is pushed tentatively with the
// current tmax; its pop tests the exact box (the child's own box, stored at
// the start of its record or leaf block) with that tmax, exactly as the
// reference tested it at the parent, and skips the visit when the test fails
// (nothing is counted, nothing pushed).  The visit sequence and the counts
// are therefore the reference's.
//
// The quantized distance of plane q on axis a is fma(q, s*inv, (F - o)*inv)
// with a margin m = ((|F| + |o|) + 256 s) * |inv| * 2^-21 + 2^-120 taken off
// the near and added to the far side.  With K = |F| + |o| + 256 s bounding
// every plane, origin and partial sum, the reference's two roundings and the
// five of this form are each within u * K * |inv| (u = 2^-24), so 8 u K |inv|
// covers them; 2^-120 covers results in the denormal range.
//
// Padded slots [c, k) hold the zero box, tested exactly as the reference does.
// Rays outside the t.fast condition push every real child tentatively.

// Exact box test of an entry's own box with the tmax of its push.  The
// fast form is computed on every lane without a branch (so the box loads
// issue with the rest of the visit's loads and are waited for once); lanes
// outside its condition redo it in the exact form.
__device__ __forceinline__ bool own_box_hit(const float4& b0, const float4& b1, const Ray& r, float tmax,
                                            bool fast) {
    Ray rr = r;
    rr.tmax = tmax;
    float t0;
    bool hit = slab_fast(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, rr, t0);
    if (!fast) hit = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, rr, t0);
    return hit;
}

// Header fields of a quantized record (header at word kQHeader).
__device__ __forceinline__ NodeHeader qload_header(const uint32_t* rec, uint32_t M, int axis) {
    NodeHeader h;
    const uint32_t* hw = rec + kQHeader;
    if (M == 4) {
        const uint4 v = ld4u(hw);
        h.innerBase = v.x & kEntryIndexMask;
        h.leafMask = v.x >> 28;
        h.leafBase = v.y & kEntryIndexMask;
        h.c = (v.y >> 28) + 1u;
        h.kcodes = v.z & 0xFFu;
        h.row = ((((uint64_t)v.w << 32) | v.z) >> (8 + 16 * axis)) & 0xFFFFu;
    } else if (M == 8) {
        const uint4 v = ld4u(hw + 4 * axis);
        h.innerBase = v.x;
        h.leafBase = v.y;
        h.leafMask = v.z & 0xFFu;
        h.c = (v.z >> 8) & 15u;
        h.kcodes = v.z >> 16;
        h.row = v.w;
    } else {
        const uint4 h0 = ld4u(hw), h1 = ld4u(hw + 4), h2 = ld4u(hw + 8);
        h.innerBase = h0.x;
        h.leafBase = h0.y;
        h.leafMask = h0.z & 0xFFFFu;
        h.c = (h0.z >> 16) & 31u;
        h.kcodes = h0.w;
        const uint32_t lo = axis == 0 ? h1.x : (axis == 1 ? h1.z : h2.x);
        const uint32_t hi = axis == 0 ? h1.y : (axis == 1 ? h1.w : h2.y);
        h.row = ((uint64_t)hi << 32) | lo;
    }
    return h;
}

// The quantized planes of a record, loaded whole: dword (a, j, hi) holds the
// lo (hi = 0) or hi bytes of axis a for slots 4j .. 4j+3.
template <int M>
struct QPlanes {
    static constexpr int kLoads = M == 4 ? 2 : (M == 8 ? 3 : 6);
    uint4 v[kLoads];
    __device__ __forceinline__ void load(const uint32_t* rec) {
        const uint32_t* p = rec + qplanes_word(M);
#pragma unroll
        for (int i = 0; i < kLoads; ++i) v[i] = ld4u(p + 4 * i);
    }
    __device__ __forceinline__ uint32_t word(int w) const {
        const uint4& q = v[w >> 2];
        const int c = w & 3;
        return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w));
    }
    __device__ __forceinline__ uint32_t dword(int a, int j, int hi) const {
        return word((int)(qplane_byte(M, a, hi) / 4) + j);
    }
};

template <int MT, int S>
__device__ __forceinline__ void trav_step_q(const SceneCtx& sc, const StackCtx& sk, Trav& t) {
    constexpr uint32_t M = (uint32_t)MT;
    const uint32_t bd = sk.stride;
    uint32_t e;
    float tp;
    if (t.hasCur) {
        e = t.cur;
        tp = t.curT;
        t.hasCur = false;
    } else {
        --t.sp;
        const uint32_t slot = (t.sp & (S - 1)) * bd;
        e = sk.stE[slot];
        tp = sk.stD[slot];
        if (t.sp >= S) {
            sk.stE[slot] = sk.spE[t.sp - S];
            sk.stD[slot] = sk.spD[t.sp - S];
        }
    }
    if (e & kEntryInvalid) {
        t.err = true;
        return;
    }
    const bool tent = (e & kEntryTent) != 0;  // pushed tentatively: test the exact box first
    const uint32_t k = ((e >> kEntryKShift) & 3u) * 4u + 4u;
    const uint32_t idx = e & kEntryIndexMaskQ;
    if (e & kEntryLeaf) {
        const float* T = sc.tris + (size_t)idx * leaf_stride(sc.Lm);
        const float4 h0 = ld4(T), h1 = ld4(T + 4);
        LeafChunk a0;
        a0.load(T, 0);
        const bool ob = own_box_hit(h0, h1, t.ray, tp, t.fast);
        if (tent && !ob) return;
        ++t.cL;
        t.cLS += k;
        const uint32_t kt = t.fast ? min(k, a0.test_slots()) : k;
        if (t.anyMode) {
            bool any = false;
            for (uint32_t s0 = 0; s0 == 0 || s0 < kt; s0 += LEAF_CHUNK) {
                LeafChunk a = a0;
                if (s0) a.load(T, s0);
#pragma unroll
                for (int j = 0; j < LEAF_CHUNK; ++j) {
                    float tt, b1, b2;
                    if (mt(a.v(0, j), a.v(1, j), a.v(2, j), a.v(3, j), a.v(4, j), a.v(5, j), a.v(6, j), a.v(7, j),
                           a.v(8, j), t.ray, tt, b1, b2))
                        any = true;
                }
            }
            if (any) {
                t.occl = true;
                t.sp = 0;
                t.hasCur = false;
            }
        } else {
            // gang-G tie rule of triIntersect, followed slot by slot (see
            // trav_step): the largest kept slot at the minimum wins
            float bt = 1000000.0f;
            uint32_t hm = 0;
            int win = -1;
            float wb1 = 0.0f, wb2 = 0.0f;
            const uint32_t inst = 0xFFFFu / ((1u << sc.G) - 1u);
            for (uint32_t s0 = 0; s0 == 0 || s0 < kt; s0 += LEAF_CHUNK) {
                LeafChunk a = a0;
                if (s0) a.load(T, s0);
#pragma unroll
                for (int j = 0; j < LEAF_CHUNK; ++j) {
                    float tt, b1, b2;
                    const bool h = mt(a.v(0, j), a.v(1, j), a.v(2, j), a.v(3, j), a.v(4, j), a.v(5, j), a.v(6, j),
                                      a.v(7, j), a.v(8, j), t.ray, tt, b1, b2);
                    const uint32_t s = s0 + (uint32_t)j;
                    const bool less = h && tt < bt;
                    const bool same = h && tt == bt && hm != 0;
                    const bool take = less || (same && (hm & (inst << (s & (sc.G - 1u)))) == 0);
                    win = take ? (int)s : win;
                    wb1 = take ? b1 : wb1;
                    wb2 = take ? b2 : wb2;
                    bt = less ? tt : bt;
                    hm = less ? (1u << s) : (same ? (hm | (1u << s)) : hm);
                }
            }
            if (hm) {
                t.ray.tmax = bt;
                t.hitEntry = e & ~kEntryTent;
                t.hitSlot = win;
                t.b1 = wb1;
                t.b2 = wb2;
            }
        }
        return;
    }
    const uint32_t* rec = sc.nodes + (size_t)idx * qrecord_words(M);
    const float4 ob0 = ld4(reinterpret_cast<const float*>(rec)), ob1 = ld4(reinterpret_cast<const float*>(rec) + 4);
    const NodeHeader hd = qload_header(rec, M, t.axis);
    QPlanes<MT> pl;
    pl.load(rec);
    const bool ownHit = own_box_hit(ob0, ob1, t.ray, tp, t.fast);
    const bool vis = !tent || ownHit;
    const uint32_t c = hd.c;
    const uint32_t realMask = (1u << c) - 1u;
    const uint32_t padMask = ((1u << k) - 1u) & ~realMask;
    // padded slots: the zero box
    bool zhit = false;
    if (padMask) {
        float z;
        zhit = t.fast ? slab_fast(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, t.ray, z)
                      : slab(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, t.ray, z);
    }
    // real slots: the quantized test
    uint32_t qh = 0;
    {
        const uint32_t ex = __float_as_uint(ob1.z);
        const float F[3] = {ob0.x, ob0.y, ob0.z};
        const float o[3] = {t.ray.ox, t.ray.oy, t.ray.oz};
        const float inv[3] = {t.ray.ix, t.ray.iy, t.ray.iz};
        float qa[3], bn[3], bf[3];
        bool pos[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float sa = __builtin_ldexpf(1.0f, (int)__builtin_amdgcn_ubfe(ex, 8 * a, 8) - kQExpBias);
            qa[a] = sa * inv[a];
            const float bb = (F[a] - o[a]) * inv[a];
            const float kk = (fabsf(F[a]) + fabsf(o[a])) + 256.0f * sa;
            const float m = __builtin_fmaf(kk * fabsf(inv[a]), 0x1p-21f, 0x1p-120f);
            bn[a] = bb - m;
            bf[a] = bb + m;
            pos[a] = inv[a] >= 0.0f;
        }
#pragma unroll
        for (int j = 0; j < (MT + 3) / 4; ++j) {
            if ((uint32_t)(4 * j) >= c) break;
            uint32_t nd[3], fd[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint32_t lo = pl.dword(a, j, 0), hi = pl.dword(a, j, 1);
                nd[a] = pos[a] ? lo : hi;
                fd[a] = pos[a] ? hi : lo;
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                float tn[3], tf[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    tn[a] = __builtin_fmaf((float)((nd[a] >> (8 * b)) & 0xFFu), qa[a], bn[a]);
                    tf[a] = __builtin_fmaf((float)((fd[a] >> (8 * b)) & 0xFFu), qa[a], bf[a]);
                }
                const float t0 = fmaxf(fmaxf(tn[0], tn[1]), fmaxf(tn[2], 0.0f));
                const float t1 = fminf(fminf(tf[0], tf[1]), fminf(tf[2], t.ray.tmax));
                qh |= (t0 <= t1 ? 1u : 0u) << (4 * j + b);
            }
        }
    }
    const uint32_t H = ((t.fast ? qh : ~0u) & realMask) | (zhit ? padMask : 0u);
    t.cN += vis ? 1u : 0u;
    t.cNS += vis ? k : 0u;
    if (!vis || H == 0) return;
    // real slots pushed because of this node's test are tentative; slots
    // whose aabbDistances entry was stale are pushed for certain
    const uint32_t tentMask = realMask & ~t.live;
    t.live |= H;
    // every entry carries the tmax of its push; only tentative ones use it
    const float tnow = t.ray.tmax;
    const uint32_t innerBase = hd.innerBase, leafBase = hd.leafBase;
    const uint32_t leafMask = hd.leafMask;
    const uint32_t kcodes = hd.kcodes;
    const uint64_t row = hd.row;
    const uint32_t innerMask = ~leafMask;
    // entry of real slot s, tentative unless its slot was stale
    auto entry_real = [&](uint32_t s) -> uint32_t {
        const bool isLeaf = __builtin_amdgcn_ubfe(leafMask, s, 1) != 0;
        const uint32_t m = isLeaf ? leafMask : innerMask;
        const uint32_t b = isLeaf ? (kEntryLeaf | leafBase) : innerBase;
        const uint32_t kc = __builtin_amdgcn_ubfe(kcodes, 2 * s, 2);
        const uint32_t tb = __builtin_amdgcn_ubfe(tentMask, s, 1) << 27;
        return (kc << kEntryKShift) | tb | (b + (uint32_t)__popc(__builtin_amdgcn_ubfe(m, 0, s)));
    };
    static_assert(kEntryTent == 1u << 27, "tentative bit");
    // push(e, s): the previous pending child goes to the stack, child s
    // becomes the pending one
    auto push = [&](uint32_t en, uint32_t s) {
        if (t.hasCur) {
            const uint32_t sl = (t.sp & (S - 1)) * bd;
            if (t.sp >= S) {
                sk.spE[t.sp - S] = sk.stE[sl];
                sk.spD[t.sp - S] = sk.stD[sl];
            }
            sk.stE[sl] = t.cur;
            sk.stD[sl] = t.curT;
            ++t.sp;
        }
        t.cur = en;
        t.curT = tnow;
        t.hasCur = true;
        t.live &= ~(1u << s);
    };
    auto fits = [&](uint32_t npush) { return npush == 0 || t.sp + npush <= sk.cap; };
    if (t.anyMode) {
        // intersectSecondary: slot order, padded slots included (their
        // FastNode id resolves through the entry table, pushed for certain)
        const uint32_t m0 = t.live;
        if (!fits((uint32_t)__popc(m0))) {
            t.err = true;
        } else {
            for (uint32_t m = m0; m; m &= m - 1u) {
                const uint32_t s = (uint32_t)__builtin_ctz(m);
                uint32_t en;
                if (s >= c) {
                    const uint32_t id = rec[7] + s;
                    en = id < sc.nodeCount ? sc.nodeEntry[id] : kEntryInvalid;
                } else {
                    en = entry_real(s);
                }
                push(en, s);
            }
        }
    } else {
        uint32_t live = t.live;
        const bool pad0 = !t.rev && c < M && (live & 1u);
        if (pad0) live &= ~1u;
        uint32_t pm = 0;
        {
            const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
#pragma unroll
            for (uint32_t q = 0; q < M; ++q) {
                const uint32_t slotq = __builtin_amdgcn_ubfe(q < 8 ? lo : hi, 4 * (q & 7u), 4);
                pm |= __builtin_amdgcn_ubfe(live, slotq, 1) << q;
            }
            pm &= realMask;
        }
        const uint32_t npush = (uint32_t)__popc(pm) + (pad0 ? 1u : 0u);
        const uint32_t rowLo = (uint32_t)row, rowHi = (uint32_t)(row >> 32);
        auto slot_at = [&](uint32_t q) -> uint32_t {
            return M <= 8 ? __builtin_amdgcn_ubfe(rowLo, 4 * q, 4)
                          : __builtin_amdgcn_ubfe(q < 8 ? rowLo : rowHi, 4 * (q & 7u), 4);
        };
        if (__ballot(t.sp + npush > (uint32_t)S) == 0) {
            // every push of the wave lands in the LDS window: each entry goes
            // straight to its stack position (see trav_step)
            if (npush) {
                const uint32_t last = t.sp + npush - 1u;
                uint32_t lastE = 0;
                if (pad0) {
                    lastE = entry_real(0);
                    sk.stE[t.sp * bd] = lastE;
                    sk.stD[t.sp * bd] = tnow;
                }
                uint32_t pos = t.rev ? t.sp : last;
                const uint32_t step = t.rev ? 1u : ~0u;
                while (pm) {
                    const uint32_t q = (uint32_t)__builtin_ctz(pm);
                    pm &= pm - 1u;
                    const uint32_t sq = slot_at(q);
                    const uint32_t en = entry_real(sq);
                    sk.stE[pos * bd] = en;
                    sk.stD[pos * bd] = tnow;
                    lastE = pos == last ? en : lastE;
                    pos += step;
                }
                t.cur = lastE;
                t.curT = tnow;
                t.hasCur = true;
                t.sp = last;
            }
            t.live &= ~realMask;
        } else if (!fits(npush)) {
            t.err = true;
        } else {
            if (pad0) push(entry_real(0), 0);
            while (pm) {
                const uint32_t q = t.rev ? (uint32_t)__builtin_ctz(pm) : 31u - (uint32_t)__builtin_clz(pm);
                pm &= ~(1u << q);
                const uint32_t sq = slot_at(q);
                push(entry_real(sq), sq);
            }
        }
    }
    if (t.err) {
        t.sp = 0;
        t.hasCur = false;
    }
}

// Writes the result of a finished ray.
// TRACKED: the step kept the hit's barycentrics (trav_step<..., TRACK>);
// otherwise they are recomputed from the hit triangle.
template <bool TRACKED = false>
__device__ __forceinline__ void trav_finish(const SceneCtx& sc, const Trav& t, rt_hit* hits, uint8_t* occluded,
                                            rt_counts* counts, uint32_t r, uint32_t* errCounter) {
    if (t.anyMode) {
        occluded[r] = t.occl ? 1 : 0;
    } else {
        rt_hit h;
        if (t.hitSlot >= 0) {
            const uint32_t ord = t.hitEntry & kEntryIndexMask;
            float b1 = t.b1, b2 = t.b2;
            if constexpr (!TRACKED) {
                const float* T = sc.tris + (size_t)ord * leaf_stride(sc.Lm);
                const uint32_t hs = (uint32_t)t.hitSlot;
                float tt;
                (void)mt(T[leaf_offset(hs, 0)], T[leaf_offset(hs, 1)], T[leaf_offset(hs, 2)], T[leaf_offset(hs, 3)],
                         T[leaf_offset(hs, 4)], T[leaf_offset(hs, 5)], T[leaf_offset(hs, 6)], T[leaf_offset(hs, 7)],
                         T[leaf_offset(hs, 8)], t.ray, tt, b1, b2);
            }
            h.tri = sc.triIds[(size_t)ord * sc.Lm + t.hitSlot];
            h.leaf = ord * 9u * sc.Lm;
            h.slot = t.hitSlot;
            h.t = t.ray.tmax;
            h.b1 = b1;
            h.b2 = b2;
        } else {
            h.tri = -1;
            h.leaf = 0;
            h.slot = -1;
            h.t = t.ray.tmax;
            h.b1 = 0.0f;
            h.b2 = 0.0f;
        }
        hits[r] = h;
    }
    if (counts) {
        rt_counts c;
        c.nodes = t.cN;
        c.leaves = t.cL;
        c.node_slots = t.cNS;
        c.leaf_slots = t.cLS;
        counts[r] = c;
    }
    if (t.err) atomicAdd(errCounter, 1u);
}

__device__ __forceinline__ SceneCtx plain_scene(const KParams& p, uint32_t MT) {
    SceneCtx sc;
    sc.nodes = p.nodes;
    sc.tris = p.tris;
    sc.triIds = p.triIds;
    sc.nodeEntry = p.nodeEntry;
    sc.rootEntry = p.rootEntry;
    sc.G = p.gang;
    sc.Lm = p.Lm;
    sc.nodeCount = p.nodeCount;
    sc.leafPrims = p.leafPrims;
    sc.boxesBounded = p.boxesBounded != 0;
    sc.M = MT;
    sc.recw = record_words(MT);
    return sc;
}

// Ray fetching.  Rays are handed out in chunks: one atomicAdd on the global
// counter reserves `chunk` consecutive rays for a wave, and the wave's idle
// lanes take them from that private range without further atomics.  (A
// single counter shared by all XCDs, hit once per refill, serializes the
// waves; chunking cuts those atomics by the chunk size.)
struct WaveQueue {
    uint32_t next = 0, end = 0;  // wave-uniform private range [next, end)
    bool exhausted = false;      // the global counter has passed n

    // Lanes with `want` get consecutive ray indices (< n); the others get n.
    __device__ __forceinline__ uint32_t fetch(uint32_t* counter, uint32_t n, bool want, uint32_t chunk) {
        const uint64_t lt = __lanemask_lt();
        uint64_t need = __ballot(want);
        uint32_t got = n;
        while (need) {
            if (next == end) {
                if (exhausted) break;
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(counter, chunk);
                base = __shfl(base, 0);
                if (base >= n) {
                    exhausted = true;
                    break;
                }
                next = base;
                end = (n - base < chunk) ? n : base + chunk;
            }
            const uint32_t take = min(end - next, (uint32_t)__popcll(need));
            const bool mine = (need >> __lane_id()) & 1u;
            const uint32_t rank = (uint32_t)__popcll(need & lt);
            const bool assigned = mine && rank < take;
            if (assigned) got = next + rank;
            need &= ~__ballot(assigned);
            next += take;
        }
        return got;
    }
    // no ray left in the private range nor behind the counter
    __device__ __forceinline__ bool drained() const { return exhausted && next == end; }
};

// Per-ray kernel.  One template body for the plain launch (one scene,
// uniform parameters) and the batched launch (BATCH: every lane carries the
// scene and item of its current ray); both are specialised on the node
// memory MT and the mode.
// Batched launch tables staged in LDS after the stack arrays (batch_table_word):
// per scene the pointers and constants a step reads, per item its rays, first
// global ray and scene (the result pointers are read from the item in device
// memory when the ray finishes).  A lane keeps only its scene and item
// indices; the step reads the scene's fields from LDS, so the per-lane scene
// state of a batched launch costs two registers instead of sixteen.
struct LdsScene {
    const uint32_t* nodes;
    const float* tris;
    const int32_t* triIds;
    const uint32_t* nodeEntry;
    uint32_t rootEntry, nodeCount, lmg, pad;  // lmg: Lm | G << 8 | boxesBounded << 16
};
struct LdsItem {
    const rt_ray* rays;
    uint32_t start, scene;
};
static_assert(sizeof(LdsScene) == 48 && sizeof(LdsItem) == 16, "table sizes the batch limits are stated in");

__host__ __device__ constexpr uint32_t batch_table_word(uint32_t S, uint32_t bd, bool q) {
    return (q ? 2u : 1u) * S * bd;
}

// Copies the batch's scene and item tables into LDS; every thread of the
// block calls it once, before its first ray.
__device__ __forceinline__ void stage_batch_tables(const KParams& p, uint32_t* smem, uint32_t tableWord) {
    LdsScene* ts = reinterpret_cast<LdsScene*>(smem + tableWord);
    LdsItem* ti = reinterpret_cast<LdsItem*>(ts + p.nScenes);
    for (uint32_t i = threadIdx.x; i < p.nScenes; i += blockDim.x) {
        const BatchScene& bs = p.scenes[i];
        LdsScene v;
        v.nodes = bs.nodes;
        v.tris = bs.tris;
        v.triIds = bs.triIds;
        v.nodeEntry = bs.nodeEntry;
        v.rootEntry = bs.rootEntry;
        v.nodeCount = bs.nodeCount;
        v.lmg = bs.Lm | bs.gang << 8 | (bs.boxesBounded ? 1u : 0u) << 16;
        v.pad = 0;
        ts[i] = v;
    }
    for (uint32_t i = threadIdx.x; i < p.nItems; i += blockDim.x) {
        const BatchItem& it = p.items[i];
        LdsItem v;
        v.rays = it.rays;
        v.start = it.start;
        v.scene = it.scene;
        ti[i] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ SceneCtx lds_scene(const LdsScene& ls, uint32_t MT) {
    SceneCtx sc;
    sc.nodes = ls.nodes;
    sc.tris = ls.tris;
    sc.triIds = ls.triIds;
    sc.nodeEntry = ls.nodeEntry;
    sc.leafPrims = nullptr;
    sc.rootEntry = ls.rootEntry;
    sc.nodeCount = ls.nodeCount;
    sc.Lm = ls.lmg & 0xFFu;
    sc.G = (ls.lmg >> 8) & 0xFFu;
    sc.boxesBounded = ((ls.lmg >> 16) & 1u) != 0;
    sc.M = MT;
    sc.recw = record_words(MT);
    return sc;
}

// Body of the per-ray kernel.  Rays come from the counter `counter`, which
// hands out indices 0 .. n-1; local index i is global ray base + i (the
// batched launch runs its groups of items through one counter each).  HIST
// adds the visit statistics.
// Q: the scene's node array holds the quantized records (trav_step_q;
// closest and any hit, the tmax of each stack entry in the distance arrays).
template <int MT, int MODE, int S, bool BATCH, bool HIST, bool Q = false>
__device__ __forceinline__ void trace_body(const KParams& p, uint32_t* counter, uint32_t base, uint32_t n) {
    extern __shared__ uint32_t smem[];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const size_t gslot = (size_t)blockIdx.x * bd + tid;
    const size_t nslots = (size_t)gridDim.x * bd;
    StackCtx sk;
    sk.stE = smem + tid;
    sk.stD = reinterpret_cast<float*>(smem + S * bd) + tid;
    sk.aabbD = reinterpret_cast<float*>(smem + 2 * S * bd) + tid;
    sk.spE = p.spill + gslot * p.spillCap;
    sk.spD = reinterpret_cast<float*>(p.spill + nslots * p.spillCap) + gslot * p.spillCap;
    sk.cap = S + p.spillCap;
    sk.stride = bd;
    sk.stDep = nullptr;
    sk.spDep = nullptr;
    sk.hist = nullptr;
    if constexpr (HIST) {
        // LDS after the stack window (and the save-distance arrays): depth
        // window, then the block's counters; spill: a third region
        const uint32_t base = (MODE == kSaveDist) ? 2 * S * bd + (uint32_t)MT * bd : S * bd;
        sk.stDep = smem + base + tid;
        sk.spDep = p.spill + 2 * nslots * p.spillCap + gslot * p.spillCap;
        sk.hist = smem + base + S * bd;
        for (uint32_t w = tid; w < kHwWords; w += bd) sk.hist[w] = 0;
        __syncthreads();
    }

    SceneCtx sc = plain_scene(p, MT);
    // batched launch: the staged tables, and the lane's scene and item
    const LdsScene* tabS = reinterpret_cast<const LdsScene*>(smem + batch_table_word(S, bd, Q));
    const LdsItem* tabI = reinterpret_cast<const LdsItem*>(tabS + p.nScenes);
    uint32_t sidx = 0, item = 0;

    bool has = false;
    WaveQueue q;
    uint32_t r = 0;
    Trav t;
    t.hasCur = false;
    t.sp = 0;
    while (true) {
        if (!q.drained()) {
            const uint32_t local = q.fetch(counter, n, !has, p.chunk);
            if (!has && local < n) {
                const uint32_t rr = base + local;
                r = rr;
                if constexpr (BATCH) {
                    // item of global ray rr: last item with start <= rr
                    uint32_t lo = 0, hi = p.nItems - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) >> 1;
                        if (tabI[mid].start <= rr) lo = mid;
                        else hi = mid - 1;
                    }
                    const LdsItem it = tabI[lo];
                    item = lo;
                    sidx = it.scene;
                    r = rr - it.start;
                    const SceneCtx scs = lds_scene(tabS[sidx], MT);
                    trav_init<Q>(t, it.rays[r], scs.rootEntry, MODE == kAny, scs.boxesBounded);
                } else {
                    trav_init<Q>(t, p.rays[r], sc.rootEntry, MODE == kAny, sc.boxesBounded);
                }
                if (MODE == kSaveDist)
                    for (int s = 0; s < MT; ++s) sk.aabbD[s * bd] = __builtin_nanf("");
                has = true;
            }
        }
        if (__ballot(has) == 0) break;
        if (!has) continue;
        if (p.sched) {
            // one phase per iteration: lanes whose next entry is of the other
            // type wait (each ray still takes its own steps in order)
            const bool leafNext = (trav_peek<S>(t, sk) & kEntryLeaf) != 0;
            const uint32_t nl = (uint32_t)__popcll(__ballot(leafNext));
            const uint32_t nn = (uint32_t)__popcll(__ballot(!leafNext));
            if (leafNext != (16u * nl >= p.sched * nn)) continue;
        }
        if constexpr (BATCH) {
            const SceneCtx scs = lds_scene(tabS[sidx], MT);
            if constexpr (Q) trav_step_q<MT, S>(scs, sk, t);
            else trav_step<MT, MODE, S, HIST, !HIST>(scs, sk, t);
            if (!trav_done(t)) continue;
            const BatchItem& it = p.items[item];
            trav_finish<!HIST>(scs, t, static_cast<rt_hit*>(it.out), static_cast<uint8_t*>(it.out), it.counts, r,
                               p.work + kErrWord);
        } else {
            if constexpr (Q) trav_step_q<MT, S>(sc, sk, t);
            else trav_step<MT, MODE, S, HIST, !HIST>(sc, sk, t);
            if (!trav_done(t)) continue;
            trav_finish<!HIST>(sc, t, p.hits, p.occluded, p.counts, r, p.work + kErrWord);
        }
        has = false;
    }
    if constexpr (HIST) {
        // flush the block's counters into the rt_trace_hist layout
        __syncthreads();
        for (uint32_t w = tid; w < kHwWords; w += bd) {
            const uint32_t v = sk.hist[w];
            if (v == 0) continue;
            uint32_t g;
            if (w < kHwLeafDepth) g = 7 + (w - kHwNodeDepth);
            else if (w < kHwChildFull) g = 7 + kHistDepthBins + (w - kHwLeafDepth);
            else if (w < kHwLeafFull) g = 7 + 2 * kHistDepthBins + (w - kHwChildFull);
            else if (w < kHwAabbTests) g = 7 + 2 * kHistDepthBins + kHistFullBins + (w - kHwLeafFull);
            else g = 3 + (w - kHwAabbTests);
            atomicAdd(reinterpret_cast<unsigned long long*>(p.hist) + g, (unsigned long long)v);
        }
    }
}

// Register budget: at most 96 VGPRs, 5 waves per SIMD (measured faster than
// the unconstrained builds at 4 waves: 100 VGPRs for the plain launch, 108
// for the batched one; 6 waves spills more and is slower).  The plain kernel
// over quantized records fits 80 VGPRs without spilling for node memories up
// to 8: 6 waves per SIMD.  The batched kernels stay at 5 waves: at 6 their
// closest-hit body spills 8 words and runs 2% slower.
template <int MT, int MODE, int S, bool Q = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Q && MT <= 8 ? 6 : 5)))
k_trace(KParams p) {
    trace_body<MT, MODE, S, false, false, Q>(p, p.work + kRayCounter, 0, p.n);
}

// Batched launch.  The host orders the items by node memory and mode into
// groups (key p.groupKey[g], global rays [p.groupEnd[g-1], p.groupEnd[g])),
// each with its own ray counter.  Every wave runs the groups in turn, each
// with the body specialised on that group's node memory and mode: it takes
// rays from group g until the group's counter is exhausted and its own rays
// are finished, then moves to group g + 1.  Lanes of one wave therefore
// always run one node memory's and one mode's code (a body that reads M per
// lane keeps every memory's registers live and spilled 59 words at 5 waves,
// a third slower; one that picks closest or any hit per lane runs a batch of
// closest-hit items 5% slower than the closest-hit body).  Only the last
// group's end is the launch's tail; a wave that reaches a drained group
// leaves it after one atomic.
template <int MT, int S, bool Q>
__device__ __forceinline__ void batch_group(const KParams& p, uint32_t key, uint32_t* counter, uint32_t base,
                                            uint32_t n) {
    if (key & 1u) trace_body<MT, kAny, S, true, false, Q>(p, counter, base, n);
    else trace_body<MT, kClosest, S, true, false, Q>(p, counter, base, n);
}

// Every scene of the batch has node memory MT.
template <int MT, int S, bool Q>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_batch(KParams p) {
    extern __shared__ uint32_t smem[];
    stage_batch_tables(p, smem, batch_table_word(S, blockDim.x, Q));
    uint32_t base = 0;
    for (uint32_t g = 0; g < p.nGroups; ++g) {
        const uint32_t end = p.groupEnd[g];
        batch_group<MT, S, Q>(p, p.groupKey[g], p.work + kGroupCounter + kGroupCounterStride * g, base, end - base);
        base = end;
    }
}

// Scenes of several node memories, each at most MMAX (8 or 16): the kernel
// holds one body per node memory up to MMAX, and its registers are those of
// the largest (without the 12- and 16-slot bodies the quantized form does
// not spill).
template <int S, bool Q, int MMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_groups(KParams p) {
    extern __shared__ uint32_t smem[];
    stage_batch_tables(p, smem, batch_table_word(S, blockDim.x, Q));
    uint32_t base = 0;
    for (uint32_t g = 0; g < p.nGroups; ++g) {
        const uint32_t end = p.groupEnd[g], key = p.groupKey[g];
        uint32_t* counter = p.work + kGroupCounter + kGroupCounterStride * g;
        switch (key >> 1) {
            case 4: batch_group<4, S, Q>(p, key, counter, base, end - base); break;
            case 8: batch_group<8, S, Q>(p, key, counter, base, end - base); break;
            case 12:
                if constexpr (MMAX > 8) batch_group<12, S, Q>(p, key, counter, base, end - base);
                break;
            case 16:
                if constexpr (MMAX > 8) batch_group<16, S, Q>(p, key, counter, base, end - base);
                break;
            default: break;
        }
        base = end;
    }
}

// Instrumented per-ray kernel (rt_trace_histograms): same traversal, plus
// visit counts by depth, child/leaf fullness and test/success counts.
template <int MT, int MODE>
__global__ void __launch_bounds__(256) k_trace_hist(KParams p) {
    trace_body<MT, MODE, kHistWindow, false, true>(p, p.work + kRayCounter, 0, p.n);
}

// Per-step workload of ray tiles (rt_trace_steps; the instrumented wide
// traversal of compactNode.cpp:358-601).  One block per tile, one ray per
// thread, blocks loop over tiles.  Each step every unfinished ray takes
// exactly one visit; the block then counts node and leaf visits, distinct
// entries (a visit is distinct if no lower thread visits the same entry in
// this step: a broadcast LDS scan) and finished rays, and thread 0 writes
// the step record.  Every loop bound is block-uniform: the tile loop, and
// the step loop, which ends when no ray of the tile is unfinished.
template <int MT, int MODE>
__global__ void __launch_bounds__(1024) k_trace_steps(KParams p) {
    extern __shared__ uint32_t smem[];
    constexpr uint32_t S = kStepWindow;
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    uint32_t* keys = smem + S * bd;
    uint32_t* ctr = keys + bd;  // node work, node unique, leaf work, leaf unique, terminations
    StackCtx sk;
    sk.stE = smem + tid;
    sk.stD = nullptr;
    sk.aabbD = nullptr;
    sk.spE = p.spill + ((size_t)blockIdx.x * bd + tid) * p.spillCap;
    sk.spD = nullptr;
    sk.cap = S + p.spillCap;
    sk.stride = bd;
    sk.stDep = nullptr;
    sk.spDep = nullptr;
    sk.hist = nullptr;
    const SceneCtx sc = plain_scene(p, MT);
    const uint32_t tiles = p.n / bd;
    const bool lane0 = __lane_id() == 0;
    // blocks may have fewer than 5 threads (tiles of 1..4 rays)
    for (uint32_t w = tid; w < 5; w += bd) ctr[w] = 0;
    __syncthreads();
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t r = tile * bd + tid;
        const rt_ray in = p.rays[r];
        Trav t;
        trav_init(t, in, sc.rootEntry, MODE == kAny, sc.boxesBounded);
        const bool start = !(MODE == kAny && __builtin_isnan(in.ox));
        if (!start) t.hasCur = false;
        uint32_t* rec = reinterpret_cast<uint32_t*>(p.steps) + (size_t)tile * p.stepCap * 5u;
        {
            const uint32_t idle = (uint32_t)__popcll(__ballot(!start));
            if (lane0) atomicAdd(&ctr[4], idle);
        }
        __syncthreads();
        if (tid == 0) {
            rec[0] = bd - ctr[4];
            rec[1] = 1;
            rec[2] = 0;
            rec[3] = 0;
            rec[4] = ctr[4];
            ctr[4] = 0;
        }
        uint32_t step = 1;
        while (__syncthreads_or(!trav_done(t))) {
            const bool active = !trav_done(t);
            const uint32_t key = active ? trav_peek<S>(t, sk) : 0xFFFFFFFFu;
            keys[tid] = key;
            if (active) trav_step<MT, MODE, S>(sc, sk, t);
            const bool fin = active && trav_done(t);
            const bool leaf = (key & kEntryLeaf) != 0;
            __syncthreads();
            bool first = active;
            for (uint32_t j = 0; j < bd; ++j) first = first && !(j < tid && keys[j] == key);
            const uint32_t nN = (uint32_t)__popcll(__ballot(active && !leaf));
            const uint32_t uN = (uint32_t)__popcll(__ballot(first && !leaf));
            const uint32_t nL = (uint32_t)__popcll(__ballot(active && leaf));
            const uint32_t uL = (uint32_t)__popcll(__ballot(first && leaf));
            const uint32_t nT = (uint32_t)__popcll(__ballot(fin));
            if (lane0) {
                atomicAdd(&ctr[0], nN);
                atomicAdd(&ctr[1], uN);
                atomicAdd(&ctr[2], nL);
                atomicAdd(&ctr[3], uL);
                atomicAdd(&ctr[4], nT);
            }
            __syncthreads();
            if (tid == 0) {
                if (step < p.stepCap)
                    for (int f = 0; f < 5; ++f) rec[(size_t)step * 5u + f] = ctr[f];
                for (int f = 0; f < 5; ++f) ctr[f] = 0;
            }
            ++step;
        }
        if (tid == 0) p.stepLen[tile] = step;
        if (p.counts) {
            rt_counts c;
            c.nodes = t.cN;
            c.leaves = t.cL;
            c.node_slots = t.cNS;
            c.leaf_slots = t.cLS;
            p.counts[r] = c;
        }
        if (t.err) atomicAdd(p.work + kErrWord, 1u);
    }
}

// Ray-stream kernel: the GPU form of intersectWideAlternative
// (fastNodeManager.h:835-1049) with continuous refill (the refill prototype,
// :277-450).  A block keeps kStreamSlots rays whose traversal state lives
// in LDS.  Each round the active rays are split into a node list and a leaf
// list by the type of their next entry; the block then runs every node
// step, then every leaf step, one ray per lane, so the lanes of a phase run
// the same code.  Each ray still takes exactly one step per round in its own
// depth-first order: results and counts equal the per-ray traversal.
constexpr uint32_t kStreamSlots = 256;

template <int MT, int MODE, int S>
__global__ void __launch_bounds__(256) k_stream(KParams p) {
    constexpr uint32_t R = kStreamSlots;
    extern __shared__ uint32_t smem[];
    // SoA state per slot
    float* sRay = reinterpret_cast<float*>(smem);                // 10 * R
    uint32_t* sCur = smem + 10 * R;
    uint32_t* sSp = sCur + R;
    uint32_t* sLive = sSp + R;
    uint32_t* sHitE = sLive + R;
    int32_t* sHitS = reinterpret_cast<int32_t*>(sHitE + R);
    uint32_t* sCnt = sHitE + 2 * R;                              // 4 * R
    uint32_t* sFlags = sCnt + 4 * R;                             // active|hasCur|occl|err|rev|axis
    uint32_t* sRid = sFlags + R;
    uint32_t* sStack = sRid + R;                                 // S * R
    uint32_t* sNodeList = sStack + S * R;
    uint32_t* sLeafList = sNodeList + R;
    uint32_t* sCount = sLeafList + R;                            // [0] nodes, [1] leaves
    const uint32_t tid = threadIdx.x;
    const size_t nslots = (size_t)gridDim.x * R;
    const SceneCtx sc = plain_scene(p, MT);

    auto stack_of = [&](uint32_t slot) {
        StackCtx k;
        k.stE = sStack + slot;
        k.stD = nullptr;
        k.aabbD = nullptr;
        k.spE = p.spill + ((size_t)blockIdx.x * R + slot) * p.spillCap;
        k.spD = nullptr;
        k.cap = S + p.spillCap;
        k.stride = R;
        (void)nslots;
        return k;
    };
    auto load = [&](uint32_t s, Trav& t) {
        t.ray.tmax = sRay[s];
        t.ray.ox = sRay[R + s];
        t.ray.oy = sRay[2 * R + s];
        t.ray.oz = sRay[3 * R + s];
        t.ray.ix = sRay[4 * R + s];
        t.ray.iy = sRay[5 * R + s];
        t.ray.iz = sRay[6 * R + s];
        t.ray.dx = sRay[7 * R + s];
        t.ray.dy = sRay[8 * R + s];
        t.ray.dz = sRay[9 * R + s];
        t.cur = sCur[s];
        t.curDist = 0.0f;
        t.sp = sSp[s];
        t.live = sLive[s];
        t.hitEntry = sHitE[s];
        t.hitSlot = sHitS[s];
        t.cN = sCnt[s];
        t.cL = sCnt[R + s];
        t.cNS = sCnt[2 * R + s];
        t.cLS = sCnt[3 * R + s];
        const uint32_t f = sFlags[s];
        t.hasCur = f & 2u;
        t.occl = f & 4u;
        t.err = f & 8u;
        t.rev = f & 16u;
        t.axis = (int)((f >> 5) & 3u);
        t.fast = f & 128u;
        t.anyMode = (MODE == kAny);
    };
    auto store = [&](uint32_t s, const Trav& t) {
        sRay[s] = t.ray.tmax;
        sCur[s] = t.cur;
        sSp[s] = t.sp;
        sLive[s] = t.live;
        sHitE[s] = t.hitEntry;
        sHitS[s] = t.hitSlot;
        sCnt[s] = t.cN;
        sCnt[R + s] = t.cL;
        sCnt[2 * R + s] = t.cNS;
        sCnt[3 * R + s] = t.cLS;
        sFlags[s] = 1u | (t.hasCur ? 2u : 0u) | (t.occl ? 4u : 0u) | (t.err ? 8u : 0u) | (t.rev ? 16u : 0u) |
                    ((uint32_t)t.axis << 5) | (t.fast ? 128u : 0u);
    };

    sFlags[tid] = 0;
    WaveQueue q;
    while (true) {
        // refill the slots this thread owns (slot == tid)
        const bool free = !(sFlags[tid] & 1u);
        if (!q.drained()) {
            const uint32_t rr = q.fetch(p.work + kRayCounter, p.n, free, p.chunk);
            if (free && rr < p.n) {
                Trav t;
                trav_init(t, p.rays[rr], sc.rootEntry, MODE == kAny, sc.boxesBounded);
                const rt_ray& in = p.rays[rr];
                sRay[R + tid] = in.ox;
                sRay[2 * R + tid] = in.oy;
                sRay[3 * R + tid] = in.oz;
                sRay[4 * R + tid] = in.ix;
                sRay[5 * R + tid] = in.iy;
                sRay[6 * R + tid] = in.iz;
                sRay[7 * R + tid] = in.dx;
                sRay[8 * R + tid] = in.dy;
                sRay[9 * R + tid] = in.dz;
                sRid[tid] = rr;
                store(tid, t);
            }
        }
        if (tid < 2) sCount[tid] = 0;
        __syncthreads();
        // classify by the type of the next entry
        const bool active = sFlags[tid] & 1u;
        bool isNode = false, isLeaf = false;
        if (active) {
            const uint32_t f = sFlags[tid];
            const bool hasCur = f & 2u;
            const uint32_t e = hasCur ? sCur[tid] : sStack[((sSp[tid] - 1) & (S - 1)) * R + tid];
            isLeaf = (e & kEntryLeaf) && !(e & kEntryInvalid);
            isNode = !isLeaf;
        }
        {
            const uint64_t mn = __ballot(isNode), ml = __ballot(isLeaf);
            const uint32_t lane = __lane_id();
            uint32_t bn = 0, bl = 0;
            if (lane == 0) {
                if (mn) bn = atomicAdd(&sCount[0], (uint32_t)__popcll(mn));
                if (ml) bl = atomicAdd(&sCount[1], (uint32_t)__popcll(ml));
            }
            bn = __shfl(bn, 0);
            bl = __shfl(bl, 0);
            const uint64_t lt = __lanemask_lt();
            if (isNode) sNodeList[bn + __popcll(mn & lt)] = tid;
            if (isLeaf) sLeafList[bl + __popcll(ml & lt)] = tid;
        }
        // both reductions are block barriers and run unconditionally; the
        // block is done when no slot is active and every wave's fetch is
        // exhausted
        const int anyActive = __syncthreads_or(active);
        const int allExhausted = __syncthreads_and(q.drained());
        if (!anyActive && allExhausted) break;
        const uint32_t nNode = sCount[0], nLeaf = sCount[1];
        // node phase
        if (tid < nNode) {
            const uint32_t s = sNodeList[tid];
            Trav t;
            load(s, t);
            trav_step<MT, MODE, S>(sc, stack_of(s), t);
            store(s, t);
        }
        __syncthreads();
        // leaf phase
        if (tid < nLeaf) {
            const uint32_t s = sLeafList[tid];
            Trav t;
            load(s, t);
            trav_step<MT, MODE, S>(sc, stack_of(s), t);
            store(s, t);
        }
        __syncthreads();
        // retire finished rays of the slot this thread owns
        if (sFlags[tid] & 1u) {
            Trav t;
            load(tid, t);
            if (trav_done(t)) {
                trav_finish(sc, t, p.hits, p.occluded, p.counts, sRid[tid], p.work + kErrWord);
                sFlags[tid] = 0;
            }
        }
        __syncthreads();
    }
}

// Cooperative kernel: W lanes (a power of two >= max(M, Lm)) trace one ray.
// Lane j tests child slot j of a node / triangle j of a leaf; a ballot gives
// the hit mask.  The pushes are placed in parallel: lane q takes position q
// of the ray's order row, and a prefix popcount of the live positions gives
// its index in the reference's push sequence (forward, backward with the
// padding first, or slot order for any-hit).  Each group's stack lives in
// LDS, `depth` entries deep, so no spill is needed.  The gang-G tie rule is a
// group min-reduction plus, per program instance (slot % G), the lowest slot
// at the minimum; the result is the highest of those.
template <int W, int MODE>
__global__ void __launch_bounds__(256) k_coop(KParams p) {
    extern __shared__ uint32_t smem[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t gl = lane & (W - 1);
    const uint32_t gbase = lane & ~(uint32_t)(W - 1);
    const uint32_t wmask = (W == 32) ? 0xFFFFFFFFu : ((1u << W) - 1u);
    const uint32_t D = p.coopDepth;
    uint32_t* stk = smem + (tid / W) * D;
    const uint32_t M = p.M, G = p.gang;
    SceneCtx sc = plain_scene(p, 0);
    sc.M = M;
    sc.recw = record_words(M);
    const uint32_t LSTRIDE = leaf_stride(sc.Lm);
    // slots of the same program instance as this lane's slot
    uint32_t resMask = 0;
    for (uint32_t j = gl % G; j < (uint32_t)W; j += G) resMask |= 1u << j;

    bool has = false;
    WaveQueue q;
    uint32_t r = 0, sp = 0, live = 0, hitEntry = 0;
    int hitSlot = -1;
    bool occl = false, err = false, rev = false, fast = false;
    int axis = 0;
    uint32_t cN = 0, cL = 0, cNS = 0, cLS = 0;
    Ray ray;
    auto group_ballot = [&](bool v) -> uint32_t { return (uint32_t)(__ballot(v) >> gbase) & wmask; };

    while (true) {
        if (!q.drained()) {
            uint32_t rr = q.fetch(p.work + kRayCounter, p.n, gl == 0 && !has, p.chunk);
            rr = __shfl(rr, (int)gbase);
            if (!has && rr < p.n) {
                r = rr;
                const rt_ray& in = p.rays[r];
                ray.tmax = in.tmax;
                ray.ox = in.ox;
                ray.oy = in.oy;
                ray.oz = in.oz;
                ray.ix = in.ix;
                ray.iy = in.iy;
                ray.iz = in.iz;
                ray.dx = in.dx;
                ray.dy = in.dy;
                ray.dz = in.dz;
                axis = max_abs_dim(ray.dx, ray.dy, ray.dz);
                const float da = axis == 0 ? ray.dx : (axis == 1 ? ray.dy : ray.dz);
                rev = da <= 0.0f;
                fast = sc.boxesBounded && ray_fast(ray);
                if (gl == 0) stk[0] = sc.rootEntry;
                sp = 1;
                live = 0;
                hitEntry = 0;
                hitSlot = -1;
                occl = false;
                err = false;
                cN = cL = cNS = cLS = 0;
                has = true;
            }
        }
        if (__ballot(has) == 0) break;
        if (!has) continue;

        --sp;
        const uint32_t e = stk[sp];
        if (e & kEntryInvalid) {
            err = true;
        } else if (e & kEntryLeaf) {
            const uint32_t k = ((e >> kEntryKShift) & 3u) * 4u + 4u;
            const float* T = sc.tris + (size_t)(e & kEntryIndexMask) * LSTRIDE;
            ++cL;
            cLS += k;
            bool hit = false;
            float t = 0.0f, b1, b2;
            if (gl < k)
                hit = mt(T[leaf_offset(gl, 0)], T[leaf_offset(gl, 1)], T[leaf_offset(gl, 2)], T[leaf_offset(gl, 3)],
                         T[leaf_offset(gl, 4)], T[leaf_offset(gl, 5)], T[leaf_offset(gl, 6)], T[leaf_offset(gl, 7)],
                         T[leaf_offset(gl, 8)], ray, t, b1, b2);
            if (MODE == kAny) {
                if (group_ballot(hit)) {
                    occl = true;
                    sp = 0;
                }
            } else {
                const bool rec = hit && t < 1000000.0f;  // tStore starts at 1e6
                float best = rec ? t : 1000000.0f;
#pragma unroll
                for (int off = W / 2; off > 0; off >>= 1) {
                    const float o = __shfl_xor(best, off);
                    best = (o < best) ? o : best;
                }
                if (best < 1000000.0f) {
                    const bool flag = rec && t == best;
                    const uint32_t F = group_ballot(flag);
                    const bool first = flag && (F & resMask & ((1u << gl) - 1u)) == 0;
                    const uint32_t Fi = group_ballot(first);
                    ray.tmax = best;
                    hitEntry = e;
                    hitSlot = 31 - __clz(Fi);
                }
            }
        } else {
            const uint32_t k = ((e >> kEntryKShift) & 3u) * 4u + 4u;
            const uint32_t* rec = sc.nodes + (size_t)(e & kEntryIndexMask) * sc.recw;
            ++cN;
            cNS += k;
            bool hit = false;
            if (gl < k) {
                const float* bp = reinterpret_cast<const float*>(rec) + gl;
                float t0;
                if (fast) hit = slab_fast(bp[0], bp[M], bp[2 * M], bp[3 * M], bp[4 * M], bp[5 * M], ray, t0);
                else hit = slab(bp[0], bp[M], bp[2 * M], bp[3 * M], bp[4 * M], bp[5 * M], ray, t0);
            }
            const uint32_t H = group_ballot(hit);
            live |= H;
            if (H) {
                const NodeHeader hd = load_header(rec, M, axis);
                const uint32_t innerBase = hd.innerBase, leafBase = hd.leafBase;
                const uint32_t leafMask = hd.leafMask, c = hd.c;
                const uint32_t kcodes = hd.kcodes;
                const uint64_t row = hd.row;
                // lane gl plays position gl of the push sequence source
                uint32_t s;
                bool valid;
                if (MODE == kAny) {
                    s = gl;
                    valid = gl < M;
                } else {
                    valid = gl < c;
                    s = valid ? (uint32_t)(row >> (4 * gl)) & 15u : 0u;
                }
                const bool liveAt = valid && ((live >> s) & 1u);
                const uint32_t P = group_ballot(liveAt);
                uint32_t idx, T;
                if (MODE == kAny || rev) {
                    idx = __popc(P & ((1u << gl) - 1u));
                    T = __popc(P);
                } else {
                    const bool pad0 = c < M && (live & 1u);
                    const uint32_t Q0 = group_ballot(valid && s == 0);
                    const uint32_t Pp = pad0 ? (P & ~Q0) : P;
                    if (pad0 && ((Q0 >> gl) & 1u)) idx = 0;
                    else idx = (pad0 ? 1u : 0u) + __popc(Pp & ~((2u << gl) - 1u));
                    T = __popc(Pp) + (pad0 ? 1u : 0u);
                }
                if (sp + T > D) {
                    err = true;
                    sp = 0;
                } else {
                    if (liveAt) {
                        uint32_t en;
                        if (s >= c) {
                            const uint32_t id = rec[6 * M + child_base_word(M)] + s;
                            en = id < sc.nodeCount ? sc.nodeEntry[id] : kEntryInvalid;
                        } else {
                            const uint32_t below = (1u << s) - 1u;
                            const uint32_t kc = ((kcodes >> (2 * s)) & 3u) << kEntryKShift;
                            en = ((leafMask >> s) & 1u) ? (kEntryLeaf | kc | (leafBase + __popc(leafMask & below)))
                                                        : (kc | (innerBase + __popc(~leafMask & below)));
                        }
                        stk[sp + idx] = en;
                    }
                    sp += T;
                    const uint32_t pushedRange = (MODE == kAny) ? M : c;
                    live &= ~((pushedRange >= 32) ? 0xFFFFFFFFu : ((1u << pushedRange) - 1u));
                }
            }
        }
        if (sp != 0) continue;
        if (gl == 0) {
            Trav t;
            t.ray = ray;
            t.hitEntry = hitEntry;
            t.hitSlot = hitSlot;
            t.cN = cN;
            t.cL = cL;
            t.cNS = cNS;
            t.cLS = cLS;
            t.occl = occl;
            t.err = err;
            t.anyMode = MODE == kAny;
            trav_finish(sc, t, p.hits, p.occluded, p.counts, r, p.work + kErrWord);
        }
        has = false;
    }
}

// Zeroes the ray counters (the plain one and the batched launch's group
// counters) before a launch.  The guard counter (its own line, kErrWord) is
// not reset: it accumulates over the launches that
// share the work block and the host reads it when it reports statistics.
__global__ void k_reset(uint32_t* work) {
    for (uint32_t i = threadIdx.x; i < kErrWord; i += blockDim.x) work[i] = 0;
}

// Self-test of recip(): for the float bit patterns first, first + 1, ...
// (count of them, wrapping at 2^32) compares recip(x) with the correctly
// rounded division 1.0f / x; a mismatch is a bit difference other than two
// NaNs.  Mismatches are counted wave-aggregated into counter[0] (64-bit) and
// the lowest mismatching pattern is kept in counter[1] (atomicMin).  values,
// if not null, receives recip(x) for each pattern.  Every thread's loop is
// bounded by count.
// Load-pipeline calibration.  Each of the first `active` lanes of every wave
// loads 16 bytes from its own 128-byte line, `iters` independent loads per
// lane; lineMask + 1 lines (a power of two, at least 64, so the lanes of one
// load touch as many different lines).  This is the shape of the per-ray
// kernel's node and leaf loads: one record per lane, a different line for
// every lane.  The buffer is zero, so the final store is never taken; it
// keeps the loads.
__global__ void __launch_bounds__(256) k_gather_probe(const uint4* __restrict__ buf, uint32_t lineMask,
                                                      uint32_t active, uint32_t iters, uint32_t* sink) {
    const uint32_t lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t acc = 0;
    if (lane < active) {
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t line = ((wave * 97u + it * 131u) * 64u + lane) & lineMask;
            const uint4 v = buf[(size_t)line * 8u];
            acc += v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    if (acc == 0x9E3779B9u) sink[lane] = acc;
}

__global__ void __launch_bounds__(256) k_check_recip(uint32_t first, uint64_t count, float* values,
                                                     unsigned long long* counter) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0, lowest = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t bits = first + (uint32_t)i;
        const float x = __uint_as_float(bits);
        const float r = recip(x);
        const float q = 1.0f / x;
        if (values) values[i] = r;
        const bool same = __float_as_uint(r) == __float_as_uint(q) || (r != r && q != q);
        if (!same) {
            ++bad;
            lowest = min(lowest, (unsigned long long)bits);
        }
    }
    // wave reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        bad += __shfl_xor(bad, off);
        lowest = min(lowest, (unsigned long long)__shfl_xor(lowest, off));
    }
    if (__lane_id() == 0 && bad) {
        atomicAdd(&counter[0], bad);
        atomicMin(&counter[1], lowest);
    }
}

__global__ void k_zero(uint32_t* p, uint32_t words) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <int M, int MODE, int S, bool Q>
void* kernel_ptr() {
    return reinterpret_cast<void*>(&k_trace<M, MODE, S, Q>);
}

template <int M, int MODE>
void* pick_stream_s(uint32_t S) {
    switch (S) {
        case 8: return reinterpret_cast<void*>(&k_stream<M, MODE, 8>);
        case 16: return reinterpret_cast<void*>(&k_stream<M, MODE, 16>);
        case 32: return reinterpret_cast<void*>(&k_stream<M, MODE, 32>);
        default: return nullptr;
    }
}

void* pick_stream(uint32_t M, int mode, uint32_t S) {
    const bool any = mode == kAny;
    switch (M) {
        case 4: return any ? pick_stream_s<4, kAny>(S) : pick_stream_s<4, kClosest>(S);
        case 8: return any ? pick_stream_s<8, kAny>(S) : pick_stream_s<8, kClosest>(S);
        case 12: return any ? pick_stream_s<12, kAny>(S) : pick_stream_s<12, kClosest>(S);
        case 16: return any ? pick_stream_s<16, kAny>(S) : pick_stream_s<16, kClosest>(S);
        default: return nullptr;
    }
}

void* pick_coop(uint32_t W, int mode) {
    const bool any = mode == kAny;
    switch (W) {
        case 4: return any ? reinterpret_cast<void*>(&k_coop<4, kAny>) : reinterpret_cast<void*>(&k_coop<4, kClosest>);
        case 8: return any ? reinterpret_cast<void*>(&k_coop<8, kAny>) : reinterpret_cast<void*>(&k_coop<8, kClosest>);
        case 16: return any ? reinterpret_cast<void*>(&k_coop<16, kAny>) : reinterpret_cast<void*>(&k_coop<16, kClosest>);
        default: return nullptr;
    }
}

size_t stream_lds(uint32_t S) { return (size_t)(10 + 11 + S + 2) * kStreamSlots * 4 + 16; }

template <int M, bool Q>
void* pick_batch_s(uint32_t S) {
    switch (S) {
        case 8: return reinterpret_cast<void*>(&k_trace_batch<M, 8, Q>);
        case 16: return reinterpret_cast<void*>(&k_trace_batch<M, 16, Q>);
        case 32: return reinterpret_cast<void*>(&k_trace_batch<M, 32, Q>);
        default: return nullptr;
    }
}

template <bool Q, int MMAX>
void* pick_groups_s(uint32_t S) {
    switch (S) {
        case 8: return reinterpret_cast<void*>(&k_trace_groups<8, Q, MMAX>);
        case 16: return reinterpret_cast<void*>(&k_trace_groups<16, Q, MMAX>);
        case 32: return reinterpret_cast<void*>(&k_trace_groups<32, Q, MMAX>);
        default: return nullptr;
    }
}

// Batched kernel: specialised on the node memory when every scene of the
// batch has the same M; otherwise the grouped kernel (kBatchMixed: node
// memories up to 16; kBatchMixedSmall: up to 8), which runs each group of
// scenes with the body specialised on its node memory.
template <bool Q>
void* pick_batch_q(uint32_t S, uint32_t M) {
    switch (M) {
        case 4: return pick_batch_s<4, Q>(S);
        case 8: return pick_batch_s<8, Q>(S);
        case 12: return pick_batch_s<12, Q>(S);
        case 16: return pick_batch_s<16, Q>(S);
        case kBatchMixed: return pick_groups_s<Q, 16>(S);
        case kBatchMixedSmall: return pick_groups_s<Q, 8>(S);
        default: return nullptr;
    }
}

void* pick_batch(uint32_t S, uint32_t M, bool q) { return q ? pick_batch_q<true>(S, M) : pick_batch_q<false>(S, M); }

template <int M, int MODE, bool Q>
void* pick_s(uint32_t S) {
    switch (S) {
        case 8: return kernel_ptr<M, MODE, 8, Q>();
        case 16: return kernel_ptr<M, MODE, 16, Q>();
        case 32: return kernel_ptr<M, MODE, 32, Q>();
        default: return nullptr;
    }
}

template <int M>
void* pick_mode(int mode, uint32_t S, bool q) {
    switch (mode) {
        case kClosest: return q ? pick_s<M, kClosest, true>(S) : pick_s<M, kClosest, false>(S);
        case kAny: return q ? pick_s<M, kAny, true>(S) : pick_s<M, kAny, false>(S);
        case kSaveDist: return q ? nullptr : pick_s<M, kSaveDist, false>(S);
        default: return nullptr;
    }
}

void* pick(uint32_t M, int mode, uint32_t S, bool q) {
    if (mode == kBatch) return pick_batch(S, M, q);
    switch (M) {
        case 4: return pick_mode<4>(mode, S, q);
        case 8: return pick_mode<8>(mode, S, q);
        case 12: return pick_mode<12>(mode, S, q);
        case 16: return pick_mode<16>(mode, S, q);
        default: return nullptr;
    }
}

// Stack window (entries), for the save-distance mode its distances and
// aabbDistances, for the quantized records the tmax of each entry.
size_t lds_bytes(uint32_t M, int mode, uint32_t S, uint32_t bd, bool q) {
    size_t words = (size_t)S * bd;
    if (mode == kSaveDist) words += (size_t)S * bd + (size_t)M * bd;
    else if (q) words += (size_t)S * bd;
    return words * 4;
}

}  // namespace

hipError_t launch_trace(uint32_t M, int mode, uint32_t S, uint32_t bd, uint32_t grid, const KParams& p,
                        hipStream_t stream, bool quantized, size_t tableBytes) {
    void* fn = pick(M, mode, S, quantized);
    if (!fn) return hipErrorInvalidValue;
    if (mode == kBatch && tableBytes < batch_table_bytes(p.nScenes, p.nItems)) return hipErrorInvalidValue;
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(bd), args, lds_bytes(M, mode, S, bd, quantized) + tableBytes,
                           stream);
}

const void* batch_kernel(uint32_t M, uint32_t S, bool quantized) { return pick(M, kBatch, S, quantized); }

const void* reset_kernel() { return reinterpret_cast<const void*>(&k_reset); }

size_t trace_lds_bytes(uint32_t M, int mode, uint32_t S, uint32_t bd, bool quantized) {
    return lds_bytes(M, mode, S, bd, quantized);
}

size_t batch_table_bytes(uint32_t nScenes, uint32_t nItems) {
    return ((size_t)nScenes * sizeof(LdsScene) + (size_t)nItems * sizeof(LdsItem) + 255) / 256 * 256;
}

namespace {

template <int M>
void* pick_hist_mode(int mode) {
    switch (mode) {
        case kClosest: return reinterpret_cast<void*>(&k_trace_hist<M, kClosest>);
        case kAny: return reinterpret_cast<void*>(&k_trace_hist<M, kAny>);
        case kSaveDist: return reinterpret_cast<void*>(&k_trace_hist<M, kSaveDist>);
        default: return nullptr;
    }
}

void* pick_hist(uint32_t M, int mode) {
    switch (M) {
        case 4: return pick_hist_mode<4>(mode);
        case 8: return pick_hist_mode<8>(mode);
        case 12: return pick_hist_mode<12>(mode);
        case 16: return pick_hist_mode<16>(mode);
        default: return nullptr;
    }
}

}  // namespace

size_t trace_hist_lds_bytes(uint32_t M, int mode, uint32_t bd) {
    return lds_bytes(M, mode, kHistWindow, bd, false) + (size_t)kHistWindow * bd * 4 + (size_t)kHwWords * 4;
}

uint32_t trace_hist_block(uint32_t M, int mode) {
    uint32_t bd = 256;
    while (bd > 64 && trace_hist_lds_bytes(M, mode, bd) > 65536) bd -= 64;
    return bd;
}

int trace_hist_occupancy(uint32_t M, int mode) {
    void* fn = pick_hist(M, mode);
    if (!fn) return 0;
    const uint32_t bd = trace_hist_block(M, mode);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), (int)bd,
                                                     trace_hist_lds_bytes(M, mode, bd)) != hipSuccess)
        return 0;
    return blocks;
}

namespace {

template <int M>
void* pick_steps_mode(int mode) {
    switch (mode) {
        case kClosest: return reinterpret_cast<void*>(&k_trace_steps<M, kClosest>);
        case kAny: return reinterpret_cast<void*>(&k_trace_steps<M, kAny>);
        default: return nullptr;
    }
}

void* pick_steps(uint32_t M, int mode) {
    switch (M) {
        case 4: return pick_steps_mode<4>(mode);
        case 8: return pick_steps_mode<8>(mode);
        case 12: return pick_steps_mode<12>(mode);
        case 16: return pick_steps_mode<16>(mode);
        default: return nullptr;
    }
}

}  // namespace

size_t trace_steps_lds_bytes(uint32_t tile) { return ((size_t)kStepWindow * tile + tile + 8) * 4; }

int trace_steps_occupancy(uint32_t M, int mode, uint32_t tile) {
    void* fn = pick_steps(M, mode);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), (int)tile,
                                                     trace_steps_lds_bytes(tile)) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_trace_steps(uint32_t M, int mode, uint32_t tile, uint32_t grid, const KParams& p,
                              hipStream_t stream) {
    void* fn = pick_steps(M, mode);
    if (!fn || !p.steps || !p.stepLen || p.stepCap == 0 || tile == 0 || tile > 1024 || p.n % tile != 0)
        return hipErrorInvalidValue;
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(tile), args, trace_steps_lds_bytes(tile), stream);
}

hipError_t launch_trace_hist(uint32_t M, int mode, uint32_t grid, const KParams& p, hipStream_t stream) {
    void* fn = pick_hist(M, mode);
    if (!fn || !p.hist || !p.leafPrims) return hipErrorInvalidValue;
    const uint32_t bd = trace_hist_block(M, mode);
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(bd), args, trace_hist_lds_bytes(M, mode, bd), stream);
}

int trace_occupancy(uint32_t M, int mode, uint32_t S, uint32_t bd, bool quantized, size_t tableBytes) {
    void* fn = pick(M, mode, S, quantized);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), (int)bd,
                                                     lds_bytes(M, mode, S, bd, quantized) + tableBytes) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_stream(uint32_t M, int mode, uint32_t S, uint32_t grid, const KParams& p, hipStream_t stream) {
    if (mode == kSaveDist) return hipErrorInvalidValue;
    void* fn = pick_stream(M, mode, S);
    if (!fn) return hipErrorInvalidValue;
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(kStreamSlots), args, stream_lds(S), stream);
}

int stream_occupancy(uint32_t M, int mode, uint32_t S) {
    void* fn = pick_stream(M, mode, S);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), (int)kStreamSlots,
                                                     stream_lds(S)) != hipSuccess)
        return 0;
    return blocks;
}

uint32_t stream_block() { return kStreamSlots; }

uint32_t coop_width(uint32_t M, uint32_t Lm) {
    const uint32_t m = M > Lm ? M : Lm;
    return m <= 4 ? 4u : (m <= 8 ? 8u : 16u);
}

hipError_t launch_coop(uint32_t W, int mode, uint32_t grid, const KParams& p, hipStream_t stream) {
    if (mode == kSaveDist || mode == kBatch) return hipErrorInvalidValue;
    void* fn = pick_coop(W, mode);
    if (!fn) return hipErrorInvalidValue;
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(256), args, (size_t)(256 / W) * p.coopDepth * 4, stream);
}

int coop_occupancy(uint32_t W, int mode, uint32_t depth) {
    void* fn = pick_coop(W, mode);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), 256,
                                                     (size_t)(256 / W) * depth * 4) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_zero(uint32_t* p, uint32_t words, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero, dim3(std::max(1u, std::min(1024u, (words + 255) / 256))), dim3(256), 0, stream, p,
                       words);
    return hipGetLastError();
}

hipError_t launch_reset(uint32_t* work, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, work);
    return hipGetLastError();
}

hipError_t launch_gather_probe(const void* buf, uint32_t lines, uint32_t active, uint32_t iters, uint32_t blocks,
                               uint32_t* sink, hipStream_t stream) {
    if (lines < 64 || (lines & (lines - 1)) || active == 0 || active > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_probe, dim3(blocks), dim3(256), 0, stream, static_cast<const uint4*>(buf), lines - 1,
                       active, iters, sink);
    return hipGetLastError();
}

hipError_t launch_check_recip(uint32_t first, uint64_t count, float* values, unsigned long long* counter,
                              hipStream_t stream) {
    const uint64_t blocks = std::min<uint64_t>(2048, std::max<uint64_t>(1, (count + 255) / 256));
    hipLaunchKernelGGL(k_check_recip, dim3((uint32_t)blocks), dim3(256), 0, stream, first, count, values, counter);
    return hipGetLastError();
}

}  // namespace rt
