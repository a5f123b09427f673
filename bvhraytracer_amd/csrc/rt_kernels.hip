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

// Moeller-Trumbore body of triIntersect / triAnyHit (rayTracerTemplate.ispc:49-75).
__device__ __forceinline__ bool mt(float p0x, float p0y, float p0z, float p1x, float p1y, float p1z,
                                   float p2x, float p2y, float p2z, const Ray& r, float& t,
                                   float& b1, float& b2) {
    const float e1x = p1x - p0x, e1y = p1y - p0y, e1z = p1z - p0z;
    const float e2x = p2x - p0x, e2y = p2y - p0y, e2z = p2z - p0z;
    const float s1x = r.dy * e2z - r.dz * e2y;
    const float s1y = r.dz * e2x - r.dx * e2z;
    const float s1z = r.dx * e2y - r.dy * e2x;
    const float divisor = s1x * e1x + s1y * e1y + s1z * e1z;
    bool hit = !(divisor == 0.0f);
    const float inv = 1.0f / divisor;
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

// maxAbsDimension (glmUtil.h:18-23)
__device__ __forceinline__ int max_abs_dim(float x, float y, float z) {
    x = fabsf(x);
    y = fabsf(y);
    z = fabsf(z);
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

// One kernel template for the plain launch (one scene, uniform parameters,
// node memory MT known at compile time) and the batched launch (BATCH: every
// lane carries the scene and item of its current ray; MT = 0 reads the node
// memory from the scene; MODE = kMixed takes closest/any per item).
template <int MT, int MODE, int S, bool BATCH>
__global__ void __launch_bounds__(256) k_trace(KParams p) {
    extern __shared__ uint32_t smem[];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const uint32_t lane = __lane_id();
    uint32_t* stE = smem + tid;                                        // stack entries
    float* stD = reinterpret_cast<float*>(smem + S * bd) + tid;        // stack distances (save dist)
    float* aabbD = reinterpret_cast<float*>(smem + 2 * S * bd) + tid;  // aabbDistances (save dist)
    const size_t gslot = (size_t)blockIdx.x * bd + tid;
    const size_t nslots = (size_t)gridDim.x * bd;
    uint32_t* spE = p.spill + gslot * p.spillCap;
    float* spD = reinterpret_cast<float*>(p.spill + nslots * p.spillCap) + gslot * p.spillCap;
    const uint32_t cap = S + p.spillCap;

    // scene / item context (uniform for the plain launch, per lane in a batch)
    const uint32_t* nodes = p.nodes;
    const float* tris = p.tris;
    const int32_t* triIds = p.triIds;
    const uint32_t* nodeEntry = p.nodeEntry;
    uint32_t rootEntry = p.rootEntry, G = p.gang, Lm = p.Lm, nodeCount = p.nodeCount;
    uint32_t Mrt = MT, recw = record_words(MT), ow = order_words(MT);
    const rt_ray* rays = p.rays;
    rt_hit* hits = p.hits;
    uint8_t* occluded = p.occluded;
    rt_counts* counts = p.counts;
    bool anyMode = (MODE == kAny);

    bool has = false, exhausted = false;
    // `cur` is the entry visited next.  The child pushed last is always the
    // next one popped, so it stays in a register instead of going through
    // the LDS stack; `sp` counts only the entries below it.
    bool hasCur = false;
    uint32_t cur = 0;
    float curDist = 0.0f;
    uint32_t r = 0, sp = 0, live = 0, hitEntry = 0;
    int hitSlot = -1;
    bool occl = false, err = false;
    uint32_t cN = 0, cL = 0, cNS = 0, cLS = 0;
    int axis = 0;
    bool rev = false;
    Ray ray;

    while (true) {
        if (!exhausted) {
            const uint64_t need = __ballot(!has);
            if (need) {
                const uint32_t cnt = (uint32_t)__popcll(need);
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)need) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(p.work, cnt);
                base = __shfl(base, (int)leader);
                if (base + cnt >= p.n) exhausted = true;
                if (!has) {
                    const uint32_t rr = base + (uint32_t)__popcll(need & __lanemask_lt());
                    if (rr < p.n) {
                        r = rr;
                        if constexpr (BATCH) {
                            // item of global ray rr: last item with start <= rr
                            uint32_t lo = 0, hi = p.nItems - 1;
                            while (lo < hi) {
                                const uint32_t mid = (lo + hi + 1) >> 1;
                                if (p.items[mid].start <= rr) lo = mid;
                                else hi = mid - 1;
                            }
                            const BatchItem& it = p.items[lo];
                            const BatchScene& sc = p.scenes[it.scene];
                            nodes = sc.nodes;
                            tris = sc.tris;
                            triIds = sc.triIds;
                            nodeEntry = sc.nodeEntry;
                            rootEntry = sc.rootEntry;
                            G = sc.gang;
                            Lm = sc.Lm;
                            nodeCount = sc.nodeCount;
                            Mrt = sc.M;
                            recw = record_words(Mrt);
                            ow = order_words(Mrt);
                            rays = it.rays;
                            hits = static_cast<rt_hit*>(it.out);
                            occluded = static_cast<uint8_t*>(it.out);
                            counts = it.counts;
                            anyMode = it.mode == kAny;
                            r = rr - it.start;
                        }
                        const rt_ray& in = rays[r];
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
                        cur = rootEntry;
                        curDist = 0.0f;
                        hasCur = true;
                        if (MODE == kSaveDist)
                            for (int s = 0; s < MT; ++s) aabbD[s * bd] = __builtin_nanf("");
                        sp = 0;
                        live = 0;
                        hitEntry = 0;
                        hitSlot = -1;
                        occl = false;
                        err = false;
                        cN = cL = cNS = cLS = 0;
                        has = true;
                    }
                }
            }
        }
        if (__ballot(has) == 0) break;
        if (!has) continue;
        const uint32_t M = BATCH ? Mrt : (uint32_t)MT;
        const uint32_t LSTRIDE = leaf_stride(Lm);

        // ---- one traversal step: take the next entry and visit it ----
        uint32_t e;
        float edist = 0.0f;
        if (hasCur) {
            e = cur;
            edist = curDist;
            hasCur = false;
        } else {
            --sp;
            const uint32_t slot = (sp & (S - 1)) * bd;
            e = stE[slot];
            if (MODE == kSaveDist) edist = stD[slot];
            if (sp >= S) {
                stE[slot] = spE[sp - S];
                if (MODE == kSaveDist) stD[slot] = spD[sp - S];
            }
        }
        bool visit = true;
        if (MODE == kSaveDist && edist > ray.tmax) visit = false;
        if (visit && (e & kEntryInvalid)) {
            err = true;
            visit = false;
        }
        if (visit) {
            const uint32_t k = ((e >> kEntryKShift) & 3u) * 4u + 4u;
            const uint32_t idx = e & kEntryIndexMask;
            if (e & kEntryLeaf) {
                const float* T = tris + (size_t)idx * LSTRIDE;
                ++cL;
                cLS += k;
                if (anyMode) {
                    bool any = false;
                    for (uint32_t s0 = 0; s0 < k; s0 += 4) {
                        const float4 a0 = ld4(T + s0), a1 = ld4(T + k + s0), a2 = ld4(T + 2 * k + s0);
                        const float4 a3 = ld4(T + 3 * k + s0), a4 = ld4(T + 4 * k + s0), a5 = ld4(T + 5 * k + s0);
                        const float4 a6 = ld4(T + 6 * k + s0), a7 = ld4(T + 7 * k + s0), a8 = ld4(T + 8 * k + s0);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            float t, b1, b2;
                            if (mt(comp(a0, j), comp(a1, j), comp(a2, j), comp(a3, j), comp(a4, j), comp(a5, j),
                                   comp(a6, j), comp(a7, j), comp(a8, j), ray, t, b1, b2))
                                any = true;
                        }
                    }
                    if (any) {
                        occl = true;
                        sp = 0;
                        hasCur = false;
                    }
                } else {
                    // gang-G tie rule of triIntersect: each program instance
                    // (slot % G) keeps its first strict minimum, the result is
                    // the largest kept slot among instances at the minimum.
                    float bt = 1000000.0f;
                    int bidx = -1;
                    uint32_t gm = (1u << G) - 1u;  // instances whose tStore == bt
                    for (uint32_t s0 = 0; s0 < k; s0 += 4) {
                        const float4 a0 = ld4(T + s0), a1 = ld4(T + k + s0), a2 = ld4(T + 2 * k + s0);
                        const float4 a3 = ld4(T + 3 * k + s0), a4 = ld4(T + 4 * k + s0), a5 = ld4(T + 5 * k + s0);
                        const float4 a6 = ld4(T + 6 * k + s0), a7 = ld4(T + 7 * k + s0), a8 = ld4(T + 8 * k + s0);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            float t, b1, b2;
                            if (mt(comp(a0, j), comp(a1, j), comp(a2, j), comp(a3, j), comp(a4, j), comp(a5, j),
                                   comp(a6, j), comp(a7, j), comp(a8, j), ray, t, b1, b2)) {
                                const uint32_t s = s0 + (uint32_t)j;
                                const uint32_t g = 1u << (s & (G - 1u));
                                if (t < bt) {
                                    bt = t;
                                    bidx = (int)s;
                                    gm = g;
                                } else if (t == bt && !(gm & g)) {
                                    bidx = (int)s;
                                    gm |= g;
                                }
                            }
                        }
                    }
                    if (bidx >= 0) {
                        ray.tmax = bt;
                        hitEntry = e;
                        hitSlot = bidx;
                    }
                }
            } else {
                const uint32_t* rec = nodes + (size_t)idx * recw;
                ++cN;
                cNS += k;
                const uint4 h0 = ld4u(rec + 6 * M);
                const uint4 h1 = ld4u(rec + 6 * M + 4);
                uint4 h2 = make_uint4(0, 0, 0, 0);
                if (ow == 2) h2 = ld4u(rec + 6 * M + 8);
                uint32_t H = 0;
                const float* bp = reinterpret_cast<const float*>(rec);
                for (uint32_t s0 = 0; s0 < k; s0 += 4) {
                    const float4 nx = ld4(bp + s0), ny = ld4(bp + M + s0), nz = ld4(bp + 2 * M + s0);
                    const float4 xx = ld4(bp + 3 * M + s0), xy = ld4(bp + 4 * M + s0), xz = ld4(bp + 5 * M + s0);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float t0;
                        if (slab(comp(nx, j), comp(ny, j), comp(nz, j), comp(xx, j), comp(xy, j), comp(xz, j), ray,
                                 t0)) {
                            H |= 1u << (s0 + (uint32_t)j);
                            if (MODE == kSaveDist) aabbD[(s0 + (uint32_t)j) * bd] = t0;
                        }
                    }
                }
                live |= H;
                if (MODE == kSaveDist || H != 0) {
                    const uint32_t innerBase = h0.x, leafBase = h0.y;
                    const uint32_t leafMask = h0.z & 0xFFFFu, c = (h0.z >> 16) & 31u;
                    const uint32_t kcodes = h0.w, childBase = h1.x;
                    uint64_t row;
                    if (ow == 1) {
                        row = axis == 0 ? h1.y : (axis == 1 ? h1.z : h1.w);
                    } else {
                        const uint32_t lo = axis == 0 ? h1.y : (axis == 1 ? h1.w : h2.y);
                        const uint32_t hi = axis == 0 ? h1.z : (axis == 1 ? h2.x : h2.z);
                        row = ((uint64_t)hi << 32) | lo;
                    }
                    auto entry_of = [&](uint32_t s) -> uint32_t {
                        if (s >= c) {  // padded slot: resolve the FastNode id like the reference
                            const uint32_t id = childBase + s;
                            return id < nodeCount ? nodeEntry[id] : kEntryInvalid;
                        }
                        const uint32_t below = (1u << s) - 1u;
                        const uint32_t kc = ((kcodes >> (2 * s)) & 3u) << kEntryKShift;
                        if ((leafMask >> s) & 1u) return kEntryLeaf | kc | (leafBase + __popc(leafMask & below));
                        return kc | (innerBase + __popc(~leafMask & below));
                    };
                    // push(s): the previous pending child goes to the stack,
                    // child s becomes the pending one (visited next)
                    auto push = [&](uint32_t s) {
                        if (hasCur) {
                            if (sp >= cap) {
                                err = true;
                                return;
                            }
                            const uint32_t sl = (sp & (S - 1)) * bd;
                            if (sp >= S) {
                                spE[sp - S] = stE[sl];
                                if (MODE == kSaveDist) spD[sp - S] = stD[sl];
                            }
                            stE[sl] = cur;
                            if (MODE == kSaveDist) stD[sl] = curDist;
                            ++sp;
                        }
                        cur = entry_of(s);
                        if (MODE == kSaveDist) {
                            curDist = aabbD[s * bd];
                            aabbD[s * bd] = __builtin_nanf("");
                        }
                        hasCur = true;
                        live &= ~(1u << s);
                    };
                    if (anyMode) {
                        // intersectSecondary: slot order 0..M-1 (fastNodeManager.h:1481-1488)
                        for (uint32_t s = 0; s < M; ++s)
                            if ((live >> s) & 1u) push(s);
                    } else if (rev) {
                        for (uint32_t q = 0; q < c; ++q) {
                            const uint32_t s = (uint32_t)(row >> (4 * q)) & 15u;
                            if ((live >> s) & 1u) push(s);
                        }
                    } else {
                        // backward over the zero-padded order row: the padding
                        // (slot 0) comes first when childCount < M
                        if (c < M && (live & 1u)) push(0);
                        for (uint32_t q = c; q-- > 0;) {
                            const uint32_t s = (uint32_t)(row >> (4 * q)) & 15u;
                            if ((live >> s) & 1u) push(s);
                        }
                    }
                    if (err) {
                        sp = 0;
                        hasCur = false;
                    }
                }
            }
        }
        if (hasCur || sp != 0) continue;

        // ---- ray finished: write its result ----
        if (anyMode) {
            occluded[r] = occl ? 1 : 0;
        } else {
            rt_hit h;
            if (hitSlot >= 0) {
                const uint32_t ord = hitEntry & kEntryIndexMask;
                const uint32_t k = ((hitEntry >> kEntryKShift) & 3u) * 4u + 4u;
                const float* T = tris + (size_t)ord * LSTRIDE + hitSlot;
                float t, b1, b2;
                (void)mt(T[0], T[k], T[2 * k], T[3 * k], T[4 * k], T[5 * k], T[6 * k], T[7 * k], T[8 * k], ray, t,
                         b1, b2);
                h.tri = triIds[(size_t)ord * Lm + hitSlot];
                h.leaf = ord * 9u * Lm;
                h.slot = hitSlot;
                h.t = ray.tmax;
                h.b1 = b1;
                h.b2 = b2;
            } else {
                h.tri = -1;
                h.leaf = 0;
                h.slot = -1;
                h.t = ray.tmax;
                h.b1 = 0.0f;
                h.b2 = 0.0f;
            }
            hits[r] = h;
        }
        if (counts) {
            rt_counts c;
            c.nodes = cN;
            c.leaves = cL;
            c.node_slots = cNS;
            c.leaf_slots = cLS;
            counts[r] = c;
        }
        if (err) atomicAdd(p.work + 1, 1u);
        has = false;
    }
}

__global__ void k_reset(uint32_t* work) {
    if (threadIdx.x < 4) work[threadIdx.x] = 0;
}

template <int M, int MODE, int S>
void* kernel_ptr() {
    return reinterpret_cast<void*>(&k_trace<M, MODE, S, false>);
}

void* pick_batch(uint32_t S) {
    switch (S) {
        case 8: return reinterpret_cast<void*>(&k_trace<0, kMixed, 8, true>);
        case 16: return reinterpret_cast<void*>(&k_trace<0, kMixed, 16, true>);
        case 32: return reinterpret_cast<void*>(&k_trace<0, kMixed, 32, true>);
        default: return nullptr;
    }
}

template <int M, int MODE>
void* pick_s(uint32_t S) {
    switch (S) {
        case 8: return kernel_ptr<M, MODE, 8>();
        case 16: return kernel_ptr<M, MODE, 16>();
        case 32: return kernel_ptr<M, MODE, 32>();
        default: return nullptr;
    }
}

template <int M>
void* pick_mode(int mode, uint32_t S) {
    switch (mode) {
        case kClosest: return pick_s<M, kClosest>(S);
        case kAny: return pick_s<M, kAny>(S);
        case kSaveDist: return pick_s<M, kSaveDist>(S);
        default: return nullptr;
    }
}

void* pick(uint32_t M, int mode, uint32_t S) {
    if (mode == kMixed) return pick_batch(S);
    switch (M) {
        case 4: return pick_mode<4>(mode, S);
        case 8: return pick_mode<8>(mode, S);
        case 12: return pick_mode<12>(mode, S);
        case 16: return pick_mode<16>(mode, S);
        default: return nullptr;
    }
}

size_t lds_bytes(uint32_t M, int mode, uint32_t S, uint32_t bd) {
    size_t words = (size_t)S * bd;
    if (mode == kSaveDist) words += (size_t)S * bd + (size_t)M * bd;
    return words * 4;
}

}  // namespace

hipError_t launch_trace(uint32_t M, int mode, uint32_t S, uint32_t bd, uint32_t grid, const KParams& p,
                        hipStream_t stream) {
    void* fn = pick(M, mode, S);
    if (!fn) return hipErrorInvalidValue;
    KParams copy = p;
    void* args[] = {&copy};
    return hipLaunchKernel(fn, dim3(grid), dim3(bd), args, lds_bytes(M, mode, S, bd), stream);
}

size_t trace_lds_bytes(uint32_t M, int mode, uint32_t S, uint32_t bd) { return lds_bytes(M, mode, S, bd); }

int trace_occupancy(uint32_t M, int mode, uint32_t S, uint32_t bd) {
    void* fn = pick(M, mode, S);
    if (!fn) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(fn), (int)bd,
                                                     lds_bytes(M, mode, S, bd)) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_reset(uint32_t* work, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, work);
    return hipGetLastError();
}

}  // namespace rt
