// Device-side C ABI: rt_upload_bvh / rt_trace / rt_stats / rt_release.
//
// rt_upload_bvh replaces FastNodeManager's role as the owner of the
// flattened BVH (fastNodeManager.h:158-215): it reads the reference FastNode
// arrays once, decodes the loop counts the reference decodes on every visit
// (getNodeLoopCount / getLeafLoopCount + callIspcTemplateNotConst,
// fastNodeManager.h:1728-1763,271-275), and writes the device layout of
// rt_layout.h into HBM.  rt_trace replaces the per-visit ISPC calls and the
// per-ray traversal loops with one launch per ray batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_common.h"
#include "rt_kernels.h"
#include "rt_layout.h"
#include "rt_shade.h"

using namespace rt;

namespace {

// Wave schedule of the per-ray kernel used unless rt_set_schedule says
// otherwise (see KParams::sched).
constexpr uint32_t kDefaultSched = 16;

int hip_fail(hipError_t e, const char* what) {
    return set_error(RT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                       \
    do {                                          \
        hipError_t _e = (expr);                   \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

// A NULL stream argument keeps the null stream's order: the library's own
// non-blocking stream first waits for the work already queued on the null
// stream (inputs made there are complete), and a call that returns before its
// work ends makes the null stream wait for it.  The library keeps its own
// stream because graph capture and per-stream scratch need a real stream.
int null_stream_before(hipStream_t own, hipEvent_t& ev) {
    if (!ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    HIP_TRY(hipEventRecord(ev, nullptr), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(own, ev, 0), "hipStreamWaitEvent");
    return RT_OK;
}

int null_stream_after(hipStream_t own, hipEvent_t ev) {
    HIP_TRY(hipEventRecord(ev, own), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(nullptr, ev, 0), "hipStreamWaitEvent");
    return RT_OK;
}

// Per-stream scratch: work counters, stack spill and host staging buffers.
struct Workspace {
    uint32_t* work = nullptr;  // kWorkWords words
    uint32_t* spill = nullptr;
    size_t spillBytes = 0;
    void* stage = nullptr;  // rays | out | counts for host-pointer calls
    size_t stageBytes = 0;
    uint64_t* hist = nullptr;  // rt_trace_hist counters of the instrumented launch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    uint64_t lastUse = 0;  // scene's call counter at the last use
    uint32_t users = 0;    // calls waiting for their launch outside the scene lock
};

// A scene keeps the scratch of at most this many streams; the least recently
// used one is freed (after its last launch) when another stream arrives.
constexpr size_t kMaxWorkspaces = 16;

void free_workspace(Workspace& w) {
    if (w.work) (void)hipFree(w.work);
    if (w.spill) (void)hipFree(w.spill);
    if (w.stage) (void)hipFree(w.stage);
    if (w.hist) (void)hipFree(w.hist);
    if (w.ev0) (void)hipEventDestroy(w.ev0);
    if (w.ev1) (void)hipEventDestroy(w.ev1);
    w = Workspace();
}

struct LayoutInfo {
    uint32_t inner = 0, leaves = 0, maxDepth = 0;
    bool boxesBounded = true;  // every stored box coordinate is within +-2^100
    bool quantized = true;     // the quantized records are valid (rt_layout.h)
};

// Frame scale of one axis of a quantized record: the smallest power of two
// s = 2^e with F + 255 s >= hi (real arithmetic), e >= -149 and e no more
// than 30 below F's exponent, so that F + q s is exact in long double for
// every q in [0, 255].
int quant_exponent(float F, float hi) {
    int eF = 0;
    (void)std::frexp(F, &eF);
    int e = std::max(-149, (F == 0.0f ? -149 : eF - 30));
    while ((long double)F + 255.0L * std::ldexp(1.0L, e) < (long double)hi) ++e;
    return e;
}

// Quantized planes of a child interval [lo, hi] in the frame (F, s = 2^e):
// the largest q_lo with F + q_lo s <= lo and the smallest q_hi with
// F + q_hi s >= hi.  False when the interval leaves [F, F + 255 s].
bool quantize(float F, int e, float lo, float hi, uint32_t& qlo, uint32_t& qhi) {
    const long double s = std::ldexp(1.0L, e), f = F;
    if ((long double)lo < f || (long double)hi > f + 255.0L * s || lo > hi) return false;
    long double a = std::floor(((long double)lo - f) / s);
    long double b = std::ceil(((long double)hi - f) / s);
    a = std::min(255.0L, std::max(0.0L, a));
    b = std::min(255.0L, std::max(0.0L, b));
    while (a > 0 && f + a * s > (long double)lo) a -= 1;
    while (a < 255 && f + (a + 1) * s <= (long double)lo) a += 1;
    while (b < 255 && f + b * s < (long double)hi) b += 1;
    while (b > 0 && f + (b - 1) * s >= (long double)hi) b -= 1;
    if (f + a * s > (long double)lo || f + b * s < (long double)hi) return false;
    qlo = (uint32_t)a;
    qhi = (uint32_t)b;
    return true;
}

uint32_t decode_node_k(const rt_bvh_desc* d, uint32_t i) {
    const uint32_t M = d->node_memory;
    int k = (int)M;
    if (M != d->gang) {
        const float* bd = d->bounds + (size_t)i * 6 * M;
        if (std::isnan(bd[M * 6 - 1])) k = (int)bd[M * 6 - 2];
    }
    return (k == 4 || k == 8 || k == 12) ? (uint32_t)k : 16u;
}

uint32_t decode_leaf_k(const rt_bvh_desc* d, uint32_t pib) {
    const uint32_t Lm = d->leaf_memory;
    int k = (int)Lm;
    if (Lm != d->gang) {
        const uint64_t end = (uint64_t)pib + (uint64_t)Lm * 9;
        if (std::isnan(d->tri_points[end - 1])) k = (int)d->tri_points[end - 2];
    }
    return (k == 4 || k == 8 || k == 12) ? (uint32_t)k : 16u;
}

}  // namespace

struct rt_scene {
    std::mutex mu;
    int device = 0;
    uint32_t M = 0, Lm = 0, G = 0, nodeCount = 0;
    uint32_t inner = 0, leaves = 0, maxDepth = 0;
    uint32_t rootEntry = 0;
    bool boxesBounded = false;
    uint32_t sched = kDefaultSched;
    uint32_t* nodes = nullptr;
    uint32_t* qnodes = nullptr;  // quantized records (null when the scene does not qualify)
    float* tris = nullptr;
    int32_t* triIds = nullptr;
    uint32_t* nodeEntry = nullptr;
    uint8_t* leafLc = nullptr;  // getLeafLoopCount per leaf block (hit reconstruction)
    uint8_t* leafPrims = nullptr;  // real triangles per leaf block (visit statistics)
    uint64_t bytes = 0;
    uint32_t blockSize = 256, stackWindow = 16;
    bool launchSet = false;  // rt_set_launch was called: its shape serves both node formats
    bool exactBoxes = false;  // rt_set_node_format: items of RT_SCENE_SHAPES batches on the exact-box records
    uint32_t cus = 0;
    std::unordered_map<hipStream_t, Workspace> ws;
    uint64_t useClock = 0;
    hipStream_t ownStream = nullptr;
    hipEvent_t joinEv = nullptr;  // orders ownStream with the null stream
    float lastMs = 0;
    uint64_t raysTraced = 0;
};

namespace {

// Builds the device layout on the host.  Validates the layout contract the
// reference's traversal relies on (contiguous children, loop counts within
// the node/leaf memory).
int convert(const rt_bvh_desc* d, std::vector<uint32_t>& nodes, std::vector<uint32_t>& qnodes,
            std::vector<float>& tris, std::vector<int32_t>& triIds, std::vector<uint32_t>& entry,
            std::vector<uint8_t>& leafLc, LayoutInfo& info) {
    const uint32_t M = d->node_memory, Lm = d->leaf_memory, G = d->gang, n = d->node_count;
    if (!(M == 4 || M == 8 || M == 12 || M == 16) || !(Lm == 4 || Lm == 8 || Lm == 12 || Lm == 16))
        return set_error(RT_ERR_UNSUPPORTED, "rt_upload_bvh: node/leaf memory must be 4, 8, 12 or 16");
    if (!(G == 4 || G == 8 || G == 16) || M % G || Lm % G)
        return set_error(RT_ERR_UNSUPPORTED, "rt_upload_bvh: gang must be 4, 8 or 16 and divide M and Lm");
    if (n == 0 || !d->bounds || !d->base || !d->order || !d->tri_points || !d->tri_ids)
        return set_error(RT_ERR_ARG, "rt_upload_bvh: missing arrays");
    const uint64_t block = 9ull * Lm;
    if (d->tri_float_count % block)
        return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: trianglePoints is not a whole number of leaf blocks");
    const uint64_t nLeafBlocks = d->tri_float_count / block;
    std::vector<uint8_t> isLeaf(n);
    std::vector<uint32_t> index(n), kOf(n);
    uint32_t nInner = 0, nLeaves = 0;
    for (uint32_t i = 0; i < n; ++i) {
        isLeaf[i] = std::isnan(d->bounds[(size_t)i * 6 * M]) ? 1 : 0;
        if (isLeaf[i]) {
            const uint32_t pib = d->base[i];
            if (pib % block || pib / block >= nLeafBlocks)
                return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: leaf primIdBegin is not a leaf block start");
            index[i] = (uint32_t)(pib / block);
            kOf[i] = decode_leaf_k(d, pib);
            if (kOf[i] > Lm) return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: leaf loop count exceeds leaf memory");
            ++nLeaves;
        } else {
            index[i] = nInner++;
            kOf[i] = decode_node_k(d, i);
            if (kOf[i] > M) return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: node loop count exceeds node memory");
        }
    }
    if (nInner >= (1u << 28) || nLeafBlocks >= (1u << 28))
        return set_error(RT_ERR_UNSUPPORTED, "rt_upload_bvh: more than 2^28 inner nodes or leaves");
    entry.resize(n);
    for (uint32_t i = 0; i < n; ++i)
        entry[i] = (isLeaf[i] ? kEntryLeaf : 0u) | (((kOf[i] / 4) - 1) << kEntryKShift) | index[i];
    const uint32_t RECW = record_words(M), OW = order_words(M);
    nodes.assign((size_t)nInner * RECW, 0u);
    // own box of every FastNode as its parent's slot holds it (the root: the
    // union of its children, filled below)
    std::vector<float> own((size_t)n * 6, 0.0f);
    std::vector<uint8_t> hasOwn(n, 0);
    std::vector<uint32_t> depth(n, 0);
    uint32_t maxDepth = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (isLeaf[i]) continue;
        uint32_t* rec = &nodes[(size_t)index[i] * RECW];
        const float* bd = d->bounds + (size_t)i * 6 * M;
        const uint32_t k = kOf[i];
        for (uint32_t p = 0; p < 6; ++p)
            for (uint32_t s = 0; s < k; ++s) {
                float v = bd[p * k + s];
                if (!(std::fabs(v) <= 0x1p100f)) info.boxesBounded = false;
                std::memcpy(&rec[p * M + s], &v, 4);
            }
        const int8_t* ord = d->order + (size_t)i * 3 * M;
        int mx = 0;
        for (uint32_t q = 0; q < M; ++q) mx = std::max<int>(mx, ord[q]);
        const uint32_t c = (uint32_t)mx + 1;
        const uint32_t cb = d->base[i];
        if ((uint64_t)cb + c > n || cb <= i)
            return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: child block outside the node array");
        for (uint32_t a = 0; a < 3; ++a) {
            std::vector<int> seen(c, 0);
            for (uint32_t q = 0; q < M; ++q) {
                const int v = ord[a * M + q];
                if (v < 0 || v >= (int)c || (q >= c && v != 0))
                    return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: traverse order is not a padded permutation");
                if (q < c) seen[v]++;
            }
            for (uint32_t s = 0; s < c; ++s)
                if (seen[s] != 1) return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: traverse order is not a permutation");
        }
        uint32_t leafMask = 0, kcodes = 0, innerBase = 0xFFFFFFFFu, leafBase = 0xFFFFFFFFu;
        uint32_t ni = 0, nl = 0;
        for (uint32_t s = 0; s < c; ++s) {
            const uint32_t ch = cb + s;
            for (uint32_t p = 0; p < 6; ++p) own[(size_t)ch * 6 + p] = bd[p * k + s];
            if (hasOwn[ch]) info.quantized = false;  // a FastNode with two parents
            hasOwn[ch] = 1;
            kcodes |= ((kOf[ch] / 4) - 1) << (2 * s);
            depth[ch] = depth[i] + 1;
            maxDepth = std::max(maxDepth, depth[ch]);
            if (isLeaf[ch]) {
                leafMask |= 1u << s;
                if (leafBase == 0xFFFFFFFFu) leafBase = index[ch];
                if (index[ch] != leafBase + nl++)
                    return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: leaf children are not contiguous");
            } else {
                if (innerBase == 0xFFFFFFFFu) innerBase = index[ch];
                if (index[ch] != innerBase + ni++)
                    return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: inner children are not contiguous");
            }
        }
        uint32_t* h = rec + 6 * M;
        const uint32_t ib = innerBase == 0xFFFFFFFFu ? 0 : innerBase;
        const uint32_t lb = leafBase == 0xFFFFFFFFu ? 0 : leafBase;
        uint64_t rows[3] = {0, 0, 0};
        for (uint32_t a = 0; a < 3; ++a)
            for (uint32_t q = 0; q < M; ++q) rows[a] |= ((uint64_t)ord[a * M + q] & 15u) << (4 * q);
        if (M <= 8 && (ib > kEntryIndexMask || lb > kEntryIndexMask))
            return set_error(RT_ERR_LAYOUT, "rt_upload_bvh: more than 2^28 inner nodes or leaves");
        if (M == 4) {
            h[0] = ib | leafMask << 28;
            h[1] = lb | (c - 1) << 28;
            const uint64_t packed = kcodes | rows[0] << 8 | rows[1] << 24 | rows[2] << 40;
            h[2] = (uint32_t)packed;
            h[3] = (uint32_t)(packed >> 32);
            h[4] = cb;
        } else if (M == 8) {
            for (uint32_t a = 0; a < 3; ++a) {
                h[4 * a] = ib;
                h[4 * a + 1] = lb;
                h[4 * a + 2] = leafMask | c << 8 | kcodes << 16;
                h[4 * a + 3] = (uint32_t)rows[a];
            }
            h[12] = cb;
        } else {
            h[0] = ib;
            h[1] = lb;
            h[2] = leafMask | (c << 16);
            h[3] = kcodes;
            h[4] = cb;
            for (uint32_t a = 0; a < 3; ++a)
                for (uint32_t w = 0; w < OW; ++w) h[5 + a * OW + w] = (uint32_t)(rows[a] >> (32 * w));
        }
    }
    if (!info.boxesBounded || nInner >= (1u << 27) || nLeafBlocks >= (1u << 27)) info.quantized = false;
    // the root's own box: the union of its children (it is never tested)
    if (!isLeaf[0] && kOf[0] > 0) {
        const float* bd = d->bounds;
        int mx = 0;
        for (uint32_t q = 0; q < M; ++q) mx = std::max<int>(mx, d->order[q]);
        for (uint32_t p = 0; p < 6; ++p) {
            float v = bd[p * kOf[0]];
            for (uint32_t s = 1; s <= (uint32_t)mx; ++s)
                v = p < 3 ? std::min(v, bd[p * kOf[0] + s]) : std::max(v, bd[p * kOf[0] + s]);
            own[p] = v;
        }
        hasOwn[0] = 1;
    }
    const uint32_t QW = qrecord_words(M), QP = qplanes_word(M);
    if (info.quantized) qnodes.assign((size_t)nInner * QW, 0u);
    for (uint32_t i = 0; i < n && info.quantized; ++i) {
        if (isLeaf[i]) continue;
        const float* ob = &own[(size_t)i * 6];
        const float* bd = d->bounds + (size_t)i * 6 * M;
        const uint32_t k = kOf[i];
        const uint32_t* rec = &nodes[(size_t)index[i] * RECW];
        uint32_t* q = &qnodes[(size_t)index[i] * QW];
        for (uint32_t p = 0; p < 6; ++p) {
            if (!(std::fabs(ob[p]) <= kQMaxCoord)) info.quantized = false;
            std::memcpy(&q[p], &ob[p], 4);
        }
        if (!info.quantized) break;
        int mx = 0;
        for (uint32_t qq = 0; qq < M; ++qq) mx = std::max<int>(mx, d->order[(size_t)i * 3 * M + qq]);
        const uint32_t c = (uint32_t)mx + 1;
        uint8_t* planes = reinterpret_cast<uint8_t*>(q + QP);
        for (uint32_t a = 0; a < 3 && info.quantized; ++a) {
            const int e = quant_exponent(ob[a], ob[3 + a]);
            q[6] |= (uint32_t)(e + kQExpBias) << (8 * a);
            for (uint32_t s = 0; s < c; ++s) {
                uint32_t qlo = 0, qhi = 0;
                const float lo = bd[a * k + s], hi = bd[(3 + a) * k + s];
                if (!(std::fabs(lo) <= kQMaxCoord && std::fabs(hi) <= kQMaxCoord) ||
                    !quantize(ob[a], e, lo, hi, qlo, qhi)) {
                    info.quantized = false;
                    break;
                }
                planes[qplane_byte(M, a, 0) + s] = (uint8_t)qlo;
                planes[qplane_byte(M, a, 1) + s] = (uint8_t)qhi;
            }
        }
        // the padded slots [c, k) must be zero boxes: the kernel tests them as such
        for (uint32_t p = 0; p < 6; ++p)
            for (uint32_t s = c; s < k; ++s)
                if (!(bd[p * k + s] == 0.0f)) info.quantized = false;
        // header: the exact record's, without childBase (w7 here)
        const uint32_t* h = rec + 6 * M;
        q[7] = h[child_base_word(M)];
        if (M == 4) {
            for (uint32_t w = 0; w < 4; ++w) q[kQHeader + w] = h[w];
        } else if (M == 8) {
            for (uint32_t w = 0; w < 12; ++w) q[kQHeader + w] = h[w];
        } else {
            for (uint32_t w = 0; w < 4; ++w) q[kQHeader + w] = h[w];
            for (uint32_t w = 0; w < 3 * OW; ++w) q[kQHeader + 4 + w] = h[5 + w];
        }
    }
    if (!info.quantized) qnodes.clear();
    // leaf blocks (rt_layout.h): the leaf's own box in the head, then slot
    // pairs of p0 as stored, e1 = p1 - p0
    // and e2 = p2 - p0 (the edge vectors triIntersect/triAnyHit form first,
    // in the same FP32 subtraction) so the kernels do not recompute them for
    // every ray
    const uint32_t LS = leaf_stride(Lm);
    tris.assign((size_t)nLeafBlocks * LS, 0.0f);
    std::vector<uint32_t> blockK(nLeafBlocks, 0);
    std::vector<uint8_t> blockSeen(nLeafBlocks, 0);
    for (uint32_t i = 0; i < n; ++i)
        if (isLeaf[i]) {
            blockK[index[i]] = kOf[i];
            // a block shared by two leaves has no single own box for its head
            if (blockSeen[index[i]]++) info.quantized = false;
            float* head = &tris[(size_t)index[i] * LS];
            for (uint32_t p = 0; p < 6; ++p) head[p] = own[(size_t)i * 6 + p];
        }
    if (!info.quantized) qnodes.clear();
    leafLc.assign(nLeafBlocks, 0);
    for (uint64_t l = 0; l < nLeafBlocks; ++l) {
        float* dst = &tris[l * LS];
        const float* src = d->tri_points + l * block;
        const uint32_t k = blockK[l];
        // slots up to the last triangle that is not all zeros, in 2-slot
        // chunks: a zero triangle (e1 = e2 = 0) has divisor 0 for every
        // finite direction and never passes the test, so rays with a finite
        // direction skip the trailing padding
        uint32_t last = 0;
        for (uint32_t j = 0; j < k; ++j)
            for (uint32_t v = 0; v < 9; ++v)
                if (!(src[v * k + j] == 0.0f)) last = j + 1;
        // the count goes into chunk 0's first padding word, so the leaf loop
        // reads it with the first pair of triangles
        const uint32_t test = (last + 1) / 2 * 2;
        std::memcpy(&dst[kLeafHead + 18], &test, sizeof test);
        // loop count as getLeafLoopCount reads it (fastNodeManager.h:1728-1746),
        // kept for hit reconstruction when it matches the layout's stride
        int lc = (int)Lm;
        if (Lm != G && std::isnan(src[block - 1])) lc = (int)src[block - 2];
        if (k != 0 && lc == (int)k) leafLc[l] = (uint8_t)lc;
        for (uint32_t c = 0; c < 3; ++c)
            for (uint32_t s = 0; s < k; ++s) {
                const float p0 = src[c * k + s];
                dst[leaf_offset(s, c)] = p0;
                dst[leaf_offset(s, 3 + c)] = src[(3 + c) * k + s] - p0;
                dst[leaf_offset(s, 6 + c)] = src[(6 + c) * k + s] - p0;
            }
    }
    triIds.assign(d->tri_ids, d->tri_ids + nLeafBlocks * Lm);
    info.inner = nInner;
    info.leaves = (uint32_t)nLeafBlocks;
    info.maxDepth = maxDepth;
    return RT_OK;
}

template <class T>
int dev_upload(T** dst, const std::vector<T>& src, uint64_t& bytes) {
    const size_t b = std::max<size_t>(src.size() * sizeof(T), 64);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), b), "hipMalloc");
    bytes += b;
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return RT_OK;
}

void free_scene(rt_scene* s) {
    if (!s) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(s->device);
    for (auto& kv : s->ws) free_workspace(kv.second);
    if (s->nodes) (void)hipFree(s->nodes);
    if (s->qnodes) (void)hipFree(s->qnodes);
    if (s->tris) (void)hipFree(s->tris);
    if (s->triIds) (void)hipFree(s->triIds);
    if (s->nodeEntry) (void)hipFree(s->nodeEntry);
    if (s->leafLc) (void)hipFree(s->leafLc);
    if (s->leafPrims) (void)hipFree(s->leafPrims);
    if (s->joinEv) (void)hipEventDestroy(s->joinEv);
    if (s->ownStream) (void)hipStreamDestroy(s->ownStream);
    (void)hipSetDevice(cur);
    delete s;
}

// Rays a wave reserves per atomicAdd: about 1/8 of an even share of the
// launch, between 64 (one per lane) and 1024.
// Rays reserved per atomicAdd: about 1/32 of a wave's share, at least 64 (a
// wave's lanes) and at most 1024.  Smaller chunks even out the end of the
// launch (a wave that reserved a big chunk of long rays finishes last):
// measured 1/8 -> 1/32 of the share raises 8M diffuse rays from 794 to 829
// Mrays/s (N8L8); the extra atomics cost nothing visible.
uint32_t ray_chunk(uint64_t n, uint64_t grid, uint32_t bd) {
    const uint64_t waves = std::max<uint64_t>(1, grid * (bd / 64));
    uint64_t c = n / (waves * 32);
    c = (c + 63) / 64 * 64;
    return (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(64, c));
}

// Stack capacity per ray: every level keeps at most M-1 pending siblings
// plus the children of the node being expanded.
uint32_t stack_capacity(const rt_scene* s) { return ((1 + (s->maxDepth + 1) * s->M + 8) + 7) / 8 * 8; }

// The per-ray closest / any-hit kernels run on the quantized records when the
// scene has them, unless the caller asks for the exact-box records.
bool use_quantized(const rt_scene* s, int mode, uint32_t flags) {
    return s->qnodes && (mode == kClosest || mode == kAny) && !(flags & RT_EXACT_BOXES);
}

// Launch shape of the quantized records until rt_set_launch is called: their
// stack carries a float beside each entry, so an 8-entry LDS window keeps the
// LDS of 6 waves per SIMD (the exact-box records default to 256 x 16).
constexpr uint32_t kQBlock = 128, kQWindow = 8;

int get_workspace(rt_scene* s, hipStream_t st, Workspace*& out) {
    auto it = s->ws.find(st);
    if (it == s->ws.end()) {
        Workspace reuse;
        if (s->ws.size() >= kMaxWorkspaces) {
            // the calls of a scene are serialised by its lock, so no other call
            // uses the evicted scratch; its last launch ends at ev1 (host-pointer
            // calls wait for their copies before returning).  Its buffers and
            // events pass to the new stream instead of being freed, so streams
            // used in rotation do not reallocate on every call.
            // scratch that a call is still waiting on is never taken; when
            // every kept scratch is busy the set grows past the cap until
            // those calls return
            auto lru = s->ws.end();
            for (auto j = s->ws.begin(); j != s->ws.end(); ++j)
                if (j->second.users == 0 && (lru == s->ws.end() || j->second.lastUse < lru->second.lastUse)) lru = j;
            if (lru != s->ws.end()) {
                if (lru->second.timed)
                    HIP_TRY(hipEventSynchronize(lru->second.ev1), "hipEventSynchronize(scratch)");
                reuse = lru->second;
                reuse.timed = false;
                s->ws.erase(lru);
            }
        }
        it = s->ws.emplace(st, reuse).first;
    }
    Workspace& w = it->second;
    w.lastUse = ++s->useClock;
    if (!w.work) {
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.work), kWorkWords * 4), "hipMalloc(work)");
        HIP_TRY(launch_zero(w.work, kWorkWords, st), "zero kernel");  // the guard counter starts at 0
        HIP_TRY(hipEventCreate(&w.ev0), "hipEventCreate");
        HIP_TRY(hipEventCreate(&w.ev1), "hipEventCreate");
    }
    out = &w;
    return RT_OK;
}

// Grows a device buffer to at least `need` bytes, by at least half its old
// size, so a sequence of slowly growing calls reallocates O(log) times.
// Callers make sure no queued work still uses the old buffer.
int ensure(void** p, size_t& have, size_t need, const char* what) {
    if (have >= need) return RT_OK;
    const size_t grow = std::max(need, have + have / 2);
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    have = 0;
    HIP_TRY(hipMalloc(p, grow), what);
    have = grow;
    return RT_OK;
}

}  // namespace

extern "C" int rt_scene_footprint(const rt_bvh_desc* d, uint64_t* bytes) {
    if (!d || !bytes) return set_error(RT_ERR_ARG, "rt_scene_footprint: null argument");
    const uint32_t M = d->node_memory, Lm = d->leaf_memory;
    if (!(M == 4 || M == 8 || M == 12 || M == 16) || !(Lm == 4 || Lm == 8 || Lm == 12 || Lm == 16))
        return set_error(RT_ERR_UNSUPPORTED, "rt_scene_footprint: node/leaf memory must be 4, 8, 12 or 16");
    uint64_t inner = 0;
    for (uint32_t i = 0; i < d->node_count; ++i)
        if (!std::isnan(d->bounds[(size_t)i * 6 * M])) ++inner;
    const uint64_t leaves = d->tri_float_count / (9ull * Lm);
    *bytes = inner * (record_words(M) + qrecord_words(M)) * 4ull + leaves * leaf_stride(Lm) * 4ull + leaves * Lm * 4ull +
             (uint64_t)d->node_count * 4ull + 2 * std::max<uint64_t>(leaves, 64);
    return RT_OK;
}

extern "C" int rt_upload_bvh(const rt_bvh_desc* d, rt_scene_t* out) {
    if (!d || !out) return set_error(RT_ERR_ARG, "rt_upload_bvh: null argument");
    std::vector<uint32_t> nodes, qnodes, entry;
    std::vector<float> tris;
    std::vector<int32_t> triIds;
    std::vector<uint8_t> leafLc, leafPrims;
    LayoutInfo info;
    int rc;
    try {
        rc = convert(d, nodes, qnodes, tris, triIds, entry, leafLc, info);
    } catch (const std::bad_alloc&) {
        rc = set_error(RT_ERR_NOMEM, "rt_upload_bvh: out of host memory");
    }
    if (rc != RT_OK) return rc;
    // real triangles per leaf block: the slots with a triangle id (padding is -1)
    leafPrims.assign(triIds.size() / d->leaf_memory, 0);
    for (size_t l = 0; l < leafPrims.size(); ++l) {
        uint32_t c = 0;
        for (uint32_t j = 0; j < d->leaf_memory; ++j) c += triIds[l * d->leaf_memory + j] >= 0 ? 1u : 0u;
        leafPrims[l] = (uint8_t)c;
    }
    rt_scene* s = new (std::nothrow) rt_scene;
    if (!s) return set_error(RT_ERR_NOMEM, "rt_upload_bvh: out of host memory");
    (void)hipGetDevice(&s->device);
    s->M = d->node_memory;
    s->Lm = d->leaf_memory;
    s->G = d->gang;
    s->nodeCount = d->node_count;
    s->inner = info.inner;
    s->leaves = info.leaves;
    s->maxDepth = info.maxDepth;
    s->boxesBounded = info.boxesBounded;
    s->rootEntry = entry[0];
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) s->cus = (uint32_t)prop.multiProcessorCount;
    if (!s->cus) s->cus = 256;
    if ((!qnodes.empty() && (rc = dev_upload(&s->qnodes, qnodes, s->bytes)) != RT_OK) ||
        (rc = dev_upload(&s->nodes, nodes, s->bytes)) != RT_OK || (rc = dev_upload(&s->tris, tris, s->bytes)) != RT_OK ||
        (rc = dev_upload(&s->triIds, triIds, s->bytes)) != RT_OK ||
        (rc = dev_upload(&s->nodeEntry, entry, s->bytes)) != RT_OK ||
        (rc = dev_upload(&s->leafLc, leafLc, s->bytes)) != RT_OK ||
        (rc = dev_upload(&s->leafPrims, leafPrims, s->bytes)) != RT_OK) {
        free_scene(s);
        return rc;
    }
    *out = s;
    return RT_OK;
}

extern "C" int rt_set_launch(rt_scene_t s, uint32_t block_size, uint32_t stack_window) {
    if (!s) return set_error(RT_ERR_ARG, "rt_set_launch: null scene");
    if (block_size && (block_size % 64 || block_size > 256))
        return set_error(RT_ERR_ARG, "rt_set_launch: block size must be 64, 128, 192 or 256");
    if (stack_window && stack_window != 8 && stack_window != 16 && stack_window != 32)
        return set_error(RT_ERR_ARG, "rt_set_launch: stack window must be 8, 16 or 32");
    std::lock_guard<std::mutex> lk(s->mu);
    if (block_size) s->blockSize = block_size;
    if (stack_window) s->stackWindow = stack_window;
    if (block_size || stack_window) s->launchSet = true;
    return RT_OK;
}

extern "C" int rt_set_node_format(rt_scene_t s, uint32_t flags) {
    if (!s) return set_error(RT_ERR_ARG, "rt_set_node_format: null scene");
    if (flags & ~RT_EXACT_BOXES) return set_error(RT_ERR_ARG, "rt_set_node_format: flags take RT_EXACT_BOXES only");
    std::lock_guard<std::mutex> lk(s->mu);
    s->exactBoxes = (flags & RT_EXACT_BOXES) != 0;
    return RT_OK;
}

extern "C" int rt_hit_surface_device(rt_scene_t s, const rt_ray* rays, const rt_hit* hits, uint32_t n,
                                     float* normal_out, float* position_out, void* hip_stream) {
    if (!s || (n && (!rays || !hits || !normal_out || !position_out)))
        return set_error(RT_ERR_ARG, "rt_hit_surface_device: bad argument");
    HIP_TRY(launch_surface(s->tris, s->leafLc, s->Lm, s->leaves, rays, hits, n, normal_out, position_out,
                           static_cast<hipStream_t>(hip_stream)),
            "surface kernel");
    return RT_OK;
}

// The stream rt_trace uses for a scene when the caller passes none.
int rt::scene_stream(rt_scene_t s, hipStream_t* out) {
    std::lock_guard<std::mutex> lk(s->mu);
    HIP_TRY(hipSetDevice(s->device), "hipSetDevice");
    if (!s->ownStream) HIP_TRY(hipStreamCreateWithFlags(&s->ownStream, hipStreamNonBlocking), "hipStreamCreate");
    const int rc = null_stream_before(s->ownStream, s->joinEv);
    if (rc != RT_OK) return rc;
    *out = s->ownStream;
    return RT_OK;
}

extern "C" int rt_set_schedule(rt_scene_t s, uint32_t leaf_ratio16) {
    if (!s) return set_error(RT_ERR_ARG, "rt_set_schedule: null scene");
    if (leaf_ratio16 > 1024) return set_error(RT_ERR_ARG, "rt_set_schedule: ratio must be at most 1024 (64.0)");
    std::lock_guard<std::mutex> lk(s->mu);
    s->sched = leaf_ratio16;
    return RT_OK;
}

namespace {
int trace_impl(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out, rt_counts* counts,
               void* hip_stream, rt_trace_hist* hist);
}

extern "C" int rt_trace(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                        rt_counts* counts, void* hip_stream) {
    return trace_impl(s, rays, n, flags, out, counts, hip_stream, nullptr);
}

extern "C" int rt_trace_histograms(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                                   rt_trace_hist* hist, void* hip_stream) {
    if (!hist) return set_error(RT_ERR_ARG, "rt_trace_histograms: null histogram");
    if (flags & (RT_STREAM | RT_COOP | RT_COUNTS | RT_NO_SYNC))
        return set_error(RT_ERR_ARG, "rt_trace_histograms: only RT_CLOSEST/RT_ANY/RT_SAVE_DIST and RT_DEVICE_PTRS");
    std::memset(hist, 0, sizeof(*hist));
    return trace_impl(s, rays, n, flags, out, nullptr, hip_stream, hist);
}

extern "C" int rt_trace_steps(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, uint32_t tile,
                              uint32_t step_cap, rt_step_stat* out, uint32_t* steps, rt_counts* counts,
                              void* hip_stream) {
    if (!s) return set_error(RT_ERR_ARG, "rt_trace_steps: null scene");
    if (flags & ~(RT_ANY | RT_DEVICE_PTRS))
        return set_error(RT_ERR_ARG, "rt_trace_steps: only RT_CLOSEST/RT_ANY and RT_DEVICE_PTRS");
    if (tile == 0 || tile > 1024) return set_error(RT_ERR_ARG, "rt_trace_steps: tile must be 1..1024 rays");
    if (n % tile != 0) return set_error(RT_ERR_ARG, "rt_trace_steps: n must be a multiple of tile");
    if (step_cap == 0) return set_error(RT_ERR_ARG, "rt_trace_steps: step_cap must be at least 1");
    if (n == 0) return RT_OK;
    if (!rays || !out || !steps) return set_error(RT_ERR_ARG, "rt_trace_steps: null rays, records or lengths");
    if (n >= 0x80000000u) return set_error(RT_ERR_ARG, "rt_trace_steps: at most 2^31-1 rays per call");
    const uint32_t tiles = n / tile;
    if ((uint64_t)tiles * step_cap * sizeof(rt_step_stat) > (1ull << 34))
        return set_error(RT_ERR_ARG, "rt_trace_steps: tiles x step_cap records exceed 16 GiB");
    const int mode = (flags & RT_ANY) ? kAny : kClosest;
    std::lock_guard<std::mutex> lk(s->mu);
    int prevDev = 0;
    (void)hipGetDevice(&prevDev);
    if (prevDev != s->device) HIP_TRY(hipSetDevice(s->device), "hipSetDevice");
    struct Restore {
        int dev, prev;
        ~Restore() {
            if (dev != prev) (void)hipSetDevice(prev);
        }
    } restore{s->device, prevDev};
    hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
    const bool joinNull = !st;  // the call waits for its work: nothing to join after
    (void)joinNull;
    if (!st) {
        if (!s->ownStream) HIP_TRY(hipStreamCreateWithFlags(&s->ownStream, hipStreamNonBlocking), "hipStreamCreate");
        st = s->ownStream;
        const int jr = null_stream_before(st, s->joinEv);
        if (jr != RT_OK) return jr;
    }
    Workspace* w = nullptr;
    int rc = get_workspace(s, st, w);
    if (rc != RT_OK) return rc;
    const int occ = std::max(1, trace_steps_occupancy(s->M, mode, tile));
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)s->cus * occ, tiles));
    const uint32_t cap = stack_capacity(s);
    const uint32_t spillCap = cap > kStepWindow ? cap - kStepWindow : 0;
    if ((rc = ensure(reinterpret_cast<void**>(&w->spill), w->spillBytes,
                     std::max<size_t>(64, (size_t)grid * tile * spillCap * 4), "hipMalloc(spill)")) != RT_OK)
        return rc;
    const size_t rayBytes = (size_t)n * sizeof(rt_ray);
    const size_t recBytes = (size_t)tiles * step_cap * sizeof(rt_step_stat);
    const size_t lenBytes = (size_t)tiles * 4;
    const size_t cntBytes = counts ? (size_t)n * sizeof(rt_counts) : 0;
    const rt_ray* dRays = rays;
    rt_step_stat* dOut = out;
    uint32_t* dLen = steps;
    rt_counts* dCounts = counts;
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    if (!dev) {
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t outOff = up(rayBytes), lenOff = outOff + up(recBytes), cntOff = lenOff + up(lenBytes);
        if ((rc = ensure(&w->stage, w->stageBytes, cntOff + cntBytes + 256, "hipMalloc(stage)")) != RT_OK) return rc;
        char* base = static_cast<char*>(w->stage);
        dRays = reinterpret_cast<const rt_ray*>(base);
        dOut = reinterpret_cast<rt_step_stat*>(base + outOff);
        dLen = reinterpret_cast<uint32_t*>(base + lenOff);
        dCounts = counts ? reinterpret_cast<rt_counts*>(base + cntOff) : nullptr;
        HIP_TRY(hipMemcpyAsync(const_cast<rt_ray*>(dRays), rays, rayBytes, hipMemcpyHostToDevice, st), "hipMemcpy(rays)");
    }
    KParams p{};
    p.nodes = s->nodes;
    p.tris = s->tris;
    p.triIds = s->triIds;
    p.nodeEntry = s->nodeEntry;
    p.rays = dRays;
    p.counts = dCounts;
    p.spill = w->spill;
    p.work = w->work;
    p.n = n;
    p.rootEntry = s->rootEntry;
    p.spillCap = spillCap;
    p.gang = s->G;
    p.Lm = s->Lm;
    p.nodeCount = s->nodeCount;
    p.M = s->M;
    p.boxesBounded = s->boxesBounded ? 1u : 0u;
    p.steps = dOut;
    p.stepLen = dLen;
    p.stepCap = step_cap;
    HIP_TRY(launch_reset(w->work, st), "reset kernel");
    HIP_TRY(launch_trace_steps(s->M, mode, tile, grid, p, st), "step statistics kernel launch");
    if (!dev) {
        HIP_TRY(hipMemcpyAsync(out, dOut, recBytes, hipMemcpyDeviceToHost, st), "hipMemcpy(records)");
        HIP_TRY(hipMemcpyAsync(steps, dLen, lenBytes, hipMemcpyDeviceToHost, st), "hipMemcpy(lengths)");
        if (cntBytes) HIP_TRY(hipMemcpyAsync(counts, dCounts, cntBytes, hipMemcpyDeviceToHost, st), "hipMemcpy(counts)");
    }
    HIP_TRY(hipStreamSynchronize(st), "step statistics kernel");
    s->raysTraced += n;
    return RT_OK;
}

namespace {
int trace_impl(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out, rt_counts* counts,
               void* hip_stream, rt_trace_hist* hist) {
    if (!s) return set_error(RT_ERR_ARG, "rt_trace: null scene");
    if (flags & ~(RT_ANY | RT_SAVE_DIST | RT_STREAM | RT_COUNTS | RT_DEVICE_PTRS | RT_NO_SYNC | RT_COOP | RT_EXACT_BOXES))
        return set_error(RT_ERR_ARG, "rt_trace: unknown flag bits");
    if (n == 0) return RT_OK;
    if (!rays || !out) return set_error(RT_ERR_ARG, "rt_trace: null rays or output");
    if ((flags & RT_COUNTS) && !counts) return set_error(RT_ERR_ARG, "rt_trace: RT_COUNTS without a counts buffer");
    if (n >= 0x80000000u) return set_error(RT_ERR_ARG, "rt_trace: at most 2^31-1 rays per call");
    if ((flags & RT_ANY) && (flags & RT_SAVE_DIST))
        return set_error(RT_ERR_ARG, "rt_trace: RT_ANY and RT_SAVE_DIST are exclusive");
    if ((flags & RT_COOP) && (flags & (RT_SAVE_DIST | RT_STREAM)))
        return set_error(RT_ERR_UNSUPPORTED, "rt_trace: the cooperative traversal has no save-distance or stream form");
    if ((flags & RT_STREAM) && (flags & RT_SAVE_DIST))
        return set_error(RT_ERR_UNSUPPORTED, "rt_trace: the ray-stream traversal has no save-distance form");
    const int mode = (flags & RT_ANY) ? kAny : ((flags & RT_SAVE_DIST) ? kSaveDist : kClosest);
    std::unique_lock<std::mutex> lk(s->mu);
    int prevDev = 0;
    (void)hipGetDevice(&prevDev);
    if (prevDev != s->device) HIP_TRY(hipSetDevice(s->device), "hipSetDevice");
    struct Restore {
        int dev, prev;
        ~Restore() {
            if (dev != prev) (void)hipSetDevice(prev);
        }
    } restore{s->device, prevDev};
    hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
    const bool joinNull = !st;
    if (!st) {
        if (!s->ownStream) HIP_TRY(hipStreamCreateWithFlags(&s->ownStream, hipStreamNonBlocking), "hipStreamCreate");
        st = s->ownStream;
        const int jr = null_stream_before(st, s->joinEv);
        if (jr != RT_OK) return jr;
    }
    Workspace* w = nullptr;
    int rc = get_workspace(s, st, w);
    if (rc != RT_OK) return rc;

    const bool stream = (flags & RT_STREAM) != 0;
    const bool coop = (flags & RT_COOP) != 0;
    const bool quant = !hist && !stream && !coop && use_quantized(s, mode, flags);
    const bool qShape = quant && !s->launchSet;
    const uint32_t S = hist ? kHistWindow : (qShape ? kQWindow : s->stackWindow);
    const uint32_t W = coop_width(s->M, s->Lm);
    const uint32_t coopDepth = (stack_capacity(s) + 63) / 64 * 64;
    if (coop && (size_t)(256 / W) * coopDepth * 4 > 65536)
        return set_error(RT_ERR_UNSUPPORTED, "rt_trace: tree too deep for the cooperative kernel's LDS stack");
    uint32_t bd = hist ? trace_hist_block(s->M, mode) : ((stream || coop) ? 256u : (qShape ? kQBlock : s->blockSize));
    while (!stream && !coop && !hist && bd > 64 && trace_lds_bytes(s->M, mode, S, bd, quant) > 65536) bd -= 64;
    const int occ = std::max(1, hist ? trace_hist_occupancy(s->M, mode)
                                     : stream ? stream_occupancy(s->M, mode, S)
                                              : (coop ? coop_occupancy(W, mode, coopDepth)
                                                      : trace_occupancy(s->M, mode, S, bd, quant)));
    const uint64_t wantBlocks = (uint64_t)s->cus * (uint64_t)occ;
    const uint64_t raysPerBlock = coop ? bd / W : bd;
    const uint64_t needBlocks = ((uint64_t)n + raysPerBlock - 1) / raysPerBlock;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min(wantBlocks, needBlocks));
    const uint32_t cap = stack_capacity(s);
    const uint32_t spillCap = cap > S ? cap - S : 0;
    const size_t slots = (size_t)grid * bd;
    const size_t spillNeed =
        std::max<size_t>(64, slots * spillCap * 4 * (hist ? 3 : ((mode == kSaveDist || quant) ? 2 : 1)));
    const bool dev = (flags & RT_DEVICE_PTRS) != 0;
    const size_t outBytes = (mode == kAny) ? (size_t)n : (size_t)n * sizeof(rt_hit);
    const size_t rayBytes = (size_t)n * sizeof(rt_ray);
    const size_t cntBytes = (flags & RT_COUNTS) ? (size_t)n * sizeof(rt_counts) : 0;
    const rt_ray* dRays = rays;
    void* dOut = out;
    rt_counts* dCounts = (flags & RT_COUNTS) ? counts : nullptr;
    const size_t outOff = (rayBytes + 255) / 256 * 256;
    const size_t cntOff = outOff + (outBytes + 255) / 256 * 256;
    const size_t stageNeed = dev ? 0 : cntOff + cntBytes + 256;
    // a call on the same stream may still wait on this scratch (outside the
    // lock): its queued work ends before a buffer is replaced
    if (w->users && (spillNeed > w->spillBytes || stageNeed > w->stageBytes))
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize(scratch)");
    if ((rc = ensure(reinterpret_cast<void**>(&w->spill), w->spillBytes, spillNeed, "hipMalloc(spill)")) != RT_OK)
        return rc;
    if (!dev) {
        if ((rc = ensure(&w->stage, w->stageBytes, stageNeed, "hipMalloc(stage)")) != RT_OK) return rc;
        char* base = static_cast<char*>(w->stage);
        dRays = reinterpret_cast<const rt_ray*>(base);
        dOut = base + outOff;
        dCounts = cntBytes ? reinterpret_cast<rt_counts*>(base + cntOff) : nullptr;
        HIP_TRY(hipMemcpyAsync(const_cast<rt_ray*>(dRays), rays, rayBytes, hipMemcpyHostToDevice, st), "hipMemcpy(rays)");
    }
    KParams p{};
    p.nodes = quant ? s->qnodes : s->nodes;
    p.tris = s->tris;
    p.triIds = s->triIds;
    p.nodeEntry = s->nodeEntry;
    p.rays = dRays;
    p.hits = mode == kAny ? nullptr : static_cast<rt_hit*>(dOut);
    p.occluded = mode == kAny ? static_cast<uint8_t*>(dOut) : nullptr;
    p.counts = dCounts;
    p.spill = w->spill;
    p.work = w->work;
    p.n = n;
    p.rootEntry = s->rootEntry;
    p.spillCap = spillCap;
    p.gang = s->G;
    p.Lm = s->Lm;
    p.nodeCount = s->nodeCount;
    p.M = s->M;
    p.coopDepth = coopDepth;
    p.chunk = ray_chunk(n, grid, bd);
    p.boxesBounded = s->boxesBounded ? 1u : 0u;
    p.sched = s->sched;
    constexpr size_t kHistWords = sizeof(rt_trace_hist) / 8;
    if (hist) {
        if (!w->hist) HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w->hist), sizeof(rt_trace_hist)), "hipMalloc(hist)");
        HIP_TRY(launch_zero(reinterpret_cast<uint32_t*>(w->hist), (uint32_t)(kHistWords * 2), st), "zero kernel");
        p.hist = w->hist;
        p.leafPrims = s->leafPrims;
    }
    HIP_TRY(launch_reset(w->work, st), "reset kernel");
    HIP_TRY(hipEventRecord(w->ev0, st), "hipEventRecord");
    if (hist) HIP_TRY(launch_trace_hist(s->M, mode, grid, p, st), "instrumented kernel launch");
    else if (coop) HIP_TRY(launch_coop(W, mode, grid, p, st), "cooperative kernel launch");
    else if (stream) HIP_TRY(launch_stream(s->M, mode, S, grid, p, st), "stream kernel launch");
    else HIP_TRY(launch_trace(s->M, mode, S, bd, grid, p, st, quant), "trace kernel launch");
    HIP_TRY(hipEventRecord(w->ev1, st), "hipEventRecord");
    w->timed = true;
    s->raysTraced += n;
    if (!dev) {
        HIP_TRY(hipMemcpyAsync(out, dOut, outBytes, hipMemcpyDeviceToHost, st), "hipMemcpy(out)");
        if (cntBytes) HIP_TRY(hipMemcpyAsync(counts, dCounts, cntBytes, hipMemcpyDeviceToHost, st), "hipMemcpy(counts)");
    }
    if (hist)
        HIP_TRY(hipMemcpyAsync(hist, w->hist, sizeof(rt_trace_hist), hipMemcpyDeviceToHost, st), "hipMemcpy(hist)");
    if (dev && (flags & RT_NO_SYNC) && joinNull) {
        const int jr = null_stream_after(st, s->joinEv);
        if (jr != RT_OK) return jr;
    }
    if (dev && (flags & RT_NO_SYNC) && !hist) return RT_OK;
    // Wait without the scene lock, so that calls on other streams enqueue and
    // run meanwhile; the busy count keeps this scratch from being handed to
    // another stream until the wait ends.
    ++w->users;
    const hipEvent_t ev0 = w->ev0, ev1 = w->ev1;
    lk.unlock();
    const hipError_t e = hipStreamSynchronize(st);
    float ms = 0;
    const bool timed = e == hipSuccess && hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess;
    lk.lock();
    --w->users;
    if (timed) s->lastMs = ms;
    if (e != hipSuccess) return hip_fail(e, "trace kernel");
    if (hist) {
        hist->rays = n;
        hist->node_visits = hist->leaf_visits = 0;
        for (int i = 0; i < RT_HIST_DEPTHS; ++i) {
            hist->node_visits += hist->node_depth[i];
            hist->leaf_visits += hist->leaf_depth[i];
        }
    }
    return RT_OK;
}
}  // namespace

// The per-ray trace into a caller-given scratch block (work counters, then
// the stack spill): no allocation, no event, no wait.  The grid is rt_trace's,
// trimmed so that every lane's spill fits in the block.
int rt::trace_in_scratch(rt_scene_t s, const rt_ray* rays, uint32_t n, uint32_t flags, void* out,
                         rt_counts* counts, hipStream_t st, uint32_t* scratchBase, size_t scratchBytes,
                         const char* who) {
    const std::string w(who);
    if (!s) return set_error(RT_ERR_ARG, w + ": null scene");
    if (flags & ~(RT_ANY | RT_SAVE_DIST | RT_COUNTS | RT_DEVICE_PTRS | RT_NO_SYNC | RT_STREAM | RT_COOP | RT_EXACT_BOXES))
        return set_error(RT_ERR_ARG, w + ": unknown flag bits");
    if (flags & (RT_STREAM | RT_COOP))
        return set_error(RT_ERR_UNSUPPORTED, w + ": per-ray traversal only (no RT_STREAM or RT_COOP)");
    if ((flags & RT_ANY) && (flags & RT_SAVE_DIST)) return set_error(RT_ERR_ARG, w + ": RT_ANY and RT_SAVE_DIST are exclusive");
    if (n == 0) return RT_OK;
    if (!rays || !out) return set_error(RT_ERR_ARG, w + ": null rays or output");
    if ((flags & RT_COUNTS) && !counts) return set_error(RT_ERR_ARG, w + ": RT_COUNTS without a counts buffer");
    if (n >= 0x80000000u) return set_error(RT_ERR_ARG, w + ": at most 2^31-1 rays per call");
    if (!scratchBase || scratchBytes < kWorkWords * 4) return set_error(RT_ERR_ARG, w + ": no scratch");
    const int mode = (flags & RT_ANY) ? kAny : ((flags & RT_SAVE_DIST) ? kSaveDist : kClosest);
    std::lock_guard<std::mutex> lk(s->mu);
    int prevDev = 0;
    (void)hipGetDevice(&prevDev);
    if (prevDev != s->device) HIP_TRY(hipSetDevice(s->device), "hipSetDevice");
    struct Restore {
        int dev, prev;
        ~Restore() {
            if (dev != prev) (void)hipSetDevice(prev);
        }
    } restore{s->device, prevDev};
    const bool quant = use_quantized(s, mode, flags);
    const bool qShape = quant && !s->launchSet;
    const uint32_t S = qShape ? kQWindow : s->stackWindow;
    uint32_t bd = qShape ? kQBlock : s->blockSize;
    while (bd > 64 && trace_lds_bytes(s->M, mode, S, bd, quant) > 65536) bd -= 64;
    const int occ = std::max(1, trace_occupancy(s->M, mode, S, bd, quant));
    const uint32_t cap = stack_capacity(s);
    const uint32_t spillCap = cap > S ? cap - S : 0;
    const size_t perBlock = (size_t)bd * spillCap * 4 * ((mode == kSaveDist || quant) ? 2 : 1);
    const size_t spillBytes = scratchBytes - kWorkWords * 4;
    if (perBlock > spillBytes)
        return set_error(RT_ERR_UNSUPPORTED, w + ": the scene's stack spill for one block of " + std::to_string(bd) +
                                                 " lanes (" + std::to_string(perBlock) +
                                                 " bytes) exceeds the scratch (" + std::to_string(spillBytes) + " bytes)");
    const uint64_t fit = perBlock ? spillBytes / perBlock : ~0ull;
    const uint64_t needBlocks = ((uint64_t)n + bd - 1) / bd;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min(std::min((uint64_t)s->cus * (uint64_t)occ, needBlocks), fit));
    KParams p{};
    p.nodes = quant ? s->qnodes : s->nodes;
    p.tris = s->tris;
    p.triIds = s->triIds;
    p.nodeEntry = s->nodeEntry;
    p.rays = rays;
    p.hits = mode == kAny ? nullptr : static_cast<rt_hit*>(out);
    p.occluded = mode == kAny ? static_cast<uint8_t*>(out) : nullptr;
    p.counts = (flags & RT_COUNTS) ? counts : nullptr;
    p.work = scratchBase;
    p.spill = scratchBase + kWorkWords;
    p.n = n;
    p.rootEntry = s->rootEntry;
    p.spillCap = spillCap;
    p.gang = s->G;
    p.Lm = s->Lm;
    p.nodeCount = s->nodeCount;
    p.M = s->M;
    p.chunk = ray_chunk(n, grid, bd);
    p.boxesBounded = s->boxesBounded ? 1u : 0u;
    p.sched = s->sched;
    HIP_TRY(launch_reset(p.work, st), "reset kernel");
    HIP_TRY(launch_trace(s->M, mode, S, bd, grid, p, st, quant), "trace kernel launch");
    s->raysTraced += n;
    return RT_OK;
}

int rt::scene_device(rt_scene_t s) { return s ? s->device : -1; }

extern "C" int rt_check_recip(uint32_t first_bits, uint64_t count, float* values, uint64_t* mismatches,
                              uint64_t* first_mismatch, void* hip_stream) {
    if (!mismatches) return set_error(RT_ERR_ARG, "rt_check_recip: null mismatch counter");
    if (count > (1ull << 32)) return set_error(RT_ERR_ARG, "rt_check_recip: at most 2^32 patterns");
    *mismatches = 0;
    if (first_mismatch) *first_mismatch = ~0ull;
    if (count == 0) return RT_OK;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&d), 2 * sizeof(unsigned long long)), "hipMalloc(counter)");
    const unsigned long long init[2] = {0ull, ~0ull};
    unsigned long long res[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(d, init, sizeof(init), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_check_recip(first_bits, count, values, d, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res, d, sizeof(res), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "rt_check_recip");
    *mismatches = res[0];
    if (first_mismatch) *first_mismatch = res[1];
    return RT_OK;
}

extern "C" int rt_measure_gather(uint32_t active_lanes, uint32_t buffer_bytes, double* wave_loads_per_s) {
    if (!wave_loads_per_s) return set_error(RT_ERR_ARG, "rt_measure_gather: null result");
    if (active_lanes == 0 || active_lanes > 64) return set_error(RT_ERR_ARG, "rt_measure_gather: active lanes 1..64");
    if (buffer_bytes < 8192 || buffer_bytes > (1u << 30) || (buffer_bytes & (buffer_bytes - 1)))
        return set_error(RT_ERR_ARG, "rt_measure_gather: buffer bytes a power of two in [8192, 2^30]");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    // 6 blocks of 4 waves per CU: 6 waves per SIMD, the per-ray kernel's
    // occupancy.  A warm-up launch of about 10 ms brings the clocks up; the
    // rate is the best of three timed launches of about 40 ms.
    const uint32_t blocks = (uint32_t)std::max(1, prop.multiProcessorCount) * 6u, iters = 65536;
    void* buf = nullptr;
    uint32_t* sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0;
    hipError_t e = hipMalloc(&buf, buffer_bytes);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sink), 64 * 4);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(buf, 0, buffer_bytes, st);
    if (e == hipSuccess) e = launch_gather_probe(buf, buffer_bytes / 128, active_lanes, iters / 4, blocks, sink, st);
    for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
        float t = 0;
        e = hipEventRecord(e0, st);
        if (e == hipSuccess) e = launch_gather_probe(buf, buffer_bytes / 128, active_lanes, iters, blocks, sink, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && (ms == 0 || t < ms)) ms = t;
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(sink);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(e, "rt_measure_gather");
    *wave_loads_per_s = ms > 0 ? (double)blocks * 4.0 * iters / (ms * 1e-3) : 0.0;
    return RT_OK;
}

extern "C" int rt_stats(rt_scene_t s, rt_stats_t* o) {
    if (!s || !o) return set_error(RT_ERR_ARG, "rt_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->ws) {
        Workspace& w = kv.second;
        if (w.timed && hipEventQuery(w.ev1) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, w.ev0, w.ev1) == hipSuccess) s->lastMs = ms;
        }
    }
    std::memset(o, 0, sizeof(*o));
    uint64_t spill = 0;
    for (auto& kv : s->ws) spill += kv.second.spillBytes + kv.second.stageBytes;
    o->bytes_resident = s->bytes + spill;
    o->inner_nodes = s->inner;
    o->leaves = s->leaves;
    o->node_record_bytes = record_words(s->M) * 4;
    o->leaf_block_bytes = leaf_stride(s->Lm) * 4;
    o->quantized_record_bytes = s->qnodes ? qrecord_words(s->M) * 4 : 0;
    o->last_kernel_ms = s->lastMs;
    o->rays_traced = s->raysTraced;
    o->block_size = s->blockSize;
    o->stack_window = s->stackWindow;
    o->max_stack = stack_capacity(s);
    // guard counters accumulate on the device, one per stream's scratch;
    // launches still in flight are waited for
    uint64_t errs = 0;
    for (auto& kv : s->ws) {
        Workspace& w = kv.second;
        if (!w.work) continue;
        if (w.timed) HIP_TRY(hipEventSynchronize(w.ev1), "hipEventSynchronize(stats)");
        uint32_t v = 0;
        HIP_TRY(hipMemcpy(&v, w.work + kErrWord, 4, hipMemcpyDeviceToHost), "hipMemcpy(errors)");
        errs += v;
    }
    o->error_rays = (uint32_t)std::min<uint64_t>(errs, 0xFFFFFFFFu);
    return RT_OK;
}

extern "C" int rt_release(rt_scene_t s) {
    if (!s) return RT_OK;
    free_scene(s);
    return RT_OK;
}

// ---------------------------------------------------------------------------
// Batched launch: many scenes and ray batches in one persistent launch.

// One persistent launch of a batched call: the items [i0, i1) of the sorted
// table, which share a node format and a launch shape.  A call without
// RT_SCENE_SHAPES is one segment; with it, one segment per distinct
// (format, block size, stack window, schedule) of the items' scenes.
// Segments are independent (each has its own work words and spill range)
// and run concurrently: independent branches of the graph, or launches on
// forked streams, so one segment's tail overlaps the others' work.
constexpr uint32_t kMaxSegs = 16;

struct BatchSeg {
    bool q = false;
    uint32_t bd = 0, S = 0, sched = 0;
    uint32_t M = 0;     // node memory of every scene of the segment, or kBatchMixed(Small)
    uint32_t maxM = 0;  // largest node memory of the segment's scenes
    uint32_t i0 = 0, i1 = 0;
    uint32_t nGroups = 0, groupKey[kMaxGroups] = {}, groupEnd[kMaxGroups] = {};
    uint64_t total = 0;
    uint32_t grid = 0, spillCap = 0;
    size_t table = 0;  // LDS bytes reserved for the staged tables
    size_t spillOff = 0;  // first word of the segment's spill range
    bool same_launch(const BatchSeg& o) const {
        return q == o.q && bd == o.bd && S == o.S && M == o.M && grid == o.grid && table == o.table &&
               spillCap == o.spillCap && spillOff == o.spillOff;
    }
};

struct rt_batch {
    std::mutex mu;
    int device = 0;
    uint32_t maxItems = 0;
    BatchItem* dItems = nullptr;
    BatchScene* dScenes = nullptr;
    BatchItem* hItems = nullptr;  // pinned staging of the tables
    BatchScene* hScenes = nullptr;
    uint32_t* work = nullptr;  // kWorkWords words per segment, kMaxSegs segments
    uint32_t* spill = nullptr;
    size_t spillBytes = 0;
    uint32_t blockSize = 256, stackWindow = 16, cus = 0;
    bool launchSet = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // direct launches of several segments: forked streams and their events
    hipEvent_t forkEv = nullptr;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> auxEv;
    bool pending = false;
    hipStream_t ownStream = nullptr;
    hipEvent_t joinEv = nullptr;  // orders ownStream with the null stream
    // captured graph: per launch segment [reset kernel] -> [batched trace kernel]
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> traceNodes;  // in launch order
    std::vector<hipKernelNodeParams> traceNodeParams;
    hipStream_t graphStream = nullptr;
    std::vector<BatchSeg> graphSegs;  // the captured segments (shape, grid, LDS tables, spill)
    uint32_t* graphSpill = nullptr;
    uint32_t maxStack = 0;  // stack entries per ray of the last launch
    float lastMs = 0;
    uint64_t raysTraced = 0;
    uint64_t graphLaunches = 0;
};

namespace {

void drop_graph(rt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->traceNodes.clear();
    b->traceNodeParams.clear();
    b->graphSegs.clear();
}

void free_batch(rt_batch* b) {
    if (!b) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(b->device);
    drop_graph(b);
    if (b->dItems) (void)hipFree(b->dItems);
    if (b->dScenes) (void)hipFree(b->dScenes);
    if (b->hItems) (void)hipHostFree(b->hItems);
    if (b->hScenes) (void)hipHostFree(b->hScenes);
    if (b->work) (void)hipFree(b->work);
    if (b->spill) (void)hipFree(b->spill);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->forkEv) (void)hipEventDestroy(b->forkEv);
    for (auto e : b->auxEv) (void)hipEventDestroy(e);
    for (auto x : b->aux) (void)hipStreamDestroy(x);
    if (b->joinEv) (void)hipEventDestroy(b->joinEv);
    if (b->ownStream) (void)hipStreamDestroy(b->ownStream);
    (void)hipSetDevice(cur);
    delete b;
}

// Builds the graph node by node: per segment a reset kernel node and the
// batched trace kernel node after it; segments are independent branches.
int build_graph(rt_batch* b, hipStream_t st, const std::vector<BatchSeg>& segs, const std::vector<KParams>& ps) {
    drop_graph(b);
    hipGraph_t gr = nullptr;
    HIP_TRY(hipGraphCreate(&gr, 0), "hipGraphCreate");
    b->graph = gr;
    for (size_t i = 0; i < segs.size(); ++i) {
        const BatchSeg& g = segs[i];
        uint32_t* work = ps[i].work;
        void* rargs[] = {&work};
        hipKernelNodeParams rp{};
        rp.func = const_cast<void*>(reset_kernel());
        rp.gridDim = dim3(1);
        rp.blockDim = dim3(64);
        rp.sharedMemBytes = 0;
        rp.kernelParams = rargs;
        rp.extra = nullptr;
        hipGraphNode_t rn = nullptr, tn = nullptr;
        HIP_TRY(hipGraphAddKernelNode(&rn, gr, nullptr, 0, &rp), "hipGraphAddKernelNode(reset)");
        KParams copy = ps[i];
        void* targs[] = {&copy};
        hipKernelNodeParams tp{};
        tp.func = const_cast<void*>(batch_kernel(g.M, g.S, g.q));
        if (!tp.func) return set_error(RT_ERR_UNSUPPORTED, "rt_batch_trace: no batched kernel for this shape");
        tp.gridDim = dim3(g.grid);
        tp.blockDim = dim3(g.bd);
        tp.sharedMemBytes = (unsigned)(trace_lds_bytes(g.M, kBatch, g.S, g.bd, g.q) + g.table);
        tp.kernelParams = targs;
        tp.extra = nullptr;
        HIP_TRY(hipGraphAddKernelNode(&tn, gr, &rn, 1, &tp), "hipGraphAddKernelNode(trace)");
        b->traceNodes.push_back(tn);
        b->traceNodeParams.push_back(tp);
    }
    HIP_TRY(hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0), "hipGraphInstantiate");
    b->graphStream = st;
    b->graphSpill = b->spill;
    b->graphSegs = segs;
    return RT_OK;
}

// Direct launches of the segments: the first on `st`, the others on forked
// streams that join `st` again.
int launch_segments(rt_batch* b, hipStream_t st, const std::vector<BatchSeg>& segs, const std::vector<KParams>& ps) {
    if (segs.size() > 1) {
        if (!b->forkEv) HIP_TRY(hipEventCreateWithFlags(&b->forkEv, hipEventDisableTiming), "hipEventCreate");
        while (b->aux.size() + 1 < segs.size()) {
            hipStream_t x = nullptr;
            hipEvent_t e = nullptr;
            HIP_TRY(hipStreamCreateWithFlags(&x, hipStreamNonBlocking), "hipStreamCreate");
            b->aux.push_back(x);
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            b->auxEv.push_back(e);
        }
        HIP_TRY(hipEventRecord(b->forkEv, st), "hipEventRecord");
    }
    for (size_t i = 0; i < segs.size(); ++i) {
        const BatchSeg& g = segs[i];
        hipStream_t x = i == 0 ? st : b->aux[i - 1];
        if (i > 0) HIP_TRY(hipStreamWaitEvent(x, b->forkEv, 0), "hipStreamWaitEvent");
        HIP_TRY(launch_reset(ps[i].work, x), "reset kernel");
        HIP_TRY(launch_trace(g.M, kBatch, g.S, g.bd, g.grid, ps[i], x, g.q, g.table), "batched trace kernel");
        if (i > 0) HIP_TRY(hipEventRecord(b->auxEv[i - 1], x), "hipEventRecord");
    }
    for (size_t i = 1; i < segs.size(); ++i) HIP_TRY(hipStreamWaitEvent(st, b->auxEv[i - 1], 0), "hipStreamWaitEvent");
    return RT_OK;
}

}  // namespace

extern "C" int rt_batch_create(uint32_t max_items, rt_batch_t* out) {
    if (!out || max_items == 0) return set_error(RT_ERR_ARG, "rt_batch_create: bad argument");
    rt_batch* b = new (std::nothrow) rt_batch;
    if (!b) return set_error(RT_ERR_NOMEM, "rt_batch_create: out of host memory");
    (void)hipGetDevice(&b->device);
    b->maxItems = max_items;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) b->cus = (uint32_t)prop.multiProcessorCount;
    if (!b->cus) b->cus = 256;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(reinterpret_cast<void**>(&b->dItems), sizeof(BatchItem) * max_items)) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void**>(&b->dScenes), sizeof(BatchScene) * max_items)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&b->hItems), sizeof(BatchItem) * max_items, hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&b->hScenes), sizeof(BatchScene) * max_items, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void**>(&b->work), kWorkWords * kMaxSegs * 4)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess ||
        // the guard counter accumulates from here
        (e = launch_zero(b->work, kWorkWords * kMaxSegs, nullptr)) != hipSuccess ||
        (e = hipStreamSynchronize(nullptr)) != hipSuccess) {
        free_batch(b);
        return hip_fail(e, "rt_batch_create");
    }
    *out = b;
    return RT_OK;
}

extern "C" int rt_batch_set_launch(rt_batch_t b, uint32_t block_size, uint32_t stack_window) {
    if (!b) return set_error(RT_ERR_ARG, "rt_batch_set_launch: null batch");
    if (block_size && (block_size % 64 || block_size > 256))
        return set_error(RT_ERR_ARG, "rt_batch_set_launch: block size must be 64, 128, 192 or 256");
    if (stack_window && stack_window != 8 && stack_window != 16 && stack_window != 32)
        return set_error(RT_ERR_ARG, "rt_batch_set_launch: stack window must be 8, 16 or 32");
    std::lock_guard<std::mutex> lk(b->mu);
    if (block_size) b->blockSize = block_size;
    if (stack_window) b->stackWindow = stack_window;
    if (block_size || stack_window) b->launchSet = true;
    return RT_OK;
}

extern "C" int rt_batch_trace(rt_batch_t b, const rt_batch_item* items, uint32_t n_items, uint32_t flags,
                              void* hip_stream) {
    if (!b) return set_error(RT_ERR_ARG, "rt_batch_trace: null batch");
    if (flags & ~(RT_GRAPH | RT_NO_SYNC | RT_EXACT_BOXES | RT_SCENE_SHAPES))
        return set_error(RT_ERR_ARG, "rt_batch_trace: unknown flag bits");
    if (n_items == 0) return RT_OK;
    if (!items) return set_error(RT_ERR_ARG, "rt_batch_trace: null items");
    if (n_items > b->maxItems) return set_error(RT_ERR_ARG, "rt_batch_trace: more items than the batch was created for");
    std::lock_guard<std::mutex> lk(b->mu);
    int prevDev = 0;
    (void)hipGetDevice(&prevDev);
    if (prevDev != b->device) HIP_TRY(hipSetDevice(b->device), "hipSetDevice");
    struct Restore {
        int dev, prev;
        ~Restore() {
            if (dev != prev) (void)hipSetDevice(prev);
        }
    } restore{b->device, prevDev};
    hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
    const bool joinNull = !st;
    if (!st) {
        if (!b->ownStream) HIP_TRY(hipStreamCreateWithFlags(&b->ownStream, hipStreamNonBlocking), "hipStreamCreate");
        st = b->ownStream;
        const int jr = null_stream_before(st, b->joinEv);
        if (jr != RT_OK) return jr;
    }
    // the staging tables are reused: the previous batch must have consumed them
    if (b->pending) {
        HIP_TRY(hipEventSynchronize(b->ev1), "previous batch");
        b->pending = false;
    }
    uint32_t nScenes = 0, used = 0;
    uint64_t total = 0;
    uint32_t cap = 0;
    std::vector<rt_scene*> scenes;
    for (uint32_t i = 0; i < n_items; ++i) {
        const rt_batch_item& it = items[i];
        if (!it.scene || (it.n && (!it.rays || !it.out)))
            return set_error(RT_ERR_ARG, "rt_batch_trace: item without scene, rays or output");
        if (it.flags & ~RT_ANY) return set_error(RT_ERR_UNSUPPORTED, "rt_batch_trace: items support RT_CLOSEST and RT_ANY only");
        if (it.scene->device != b->device) return set_error(RT_ERR_ARG, "rt_batch_trace: scene on another device");
        if (it.n == 0) continue;
        uint32_t si = 0;
        while (si < nScenes && scenes[si] != it.scene) ++si;
        if (si == nScenes) {
            rt_scene* s = it.scene;
            scenes.push_back(s);
            BatchScene& bs = b->hScenes[nScenes++];
            bs.nodes = s->nodes;  // the quantized records replace these below when every scene has them
            bs.tris = s->tris;
            bs.triIds = s->triIds;
            bs.nodeEntry = s->nodeEntry;
            bs.rootEntry = s->rootEntry;
            bs.gang = s->G;
            bs.Lm = s->Lm;
            bs.nodeCount = s->nodeCount;
            bs.M = s->M;
            bs.boxesBounded = s->boxesBounded ? 1u : 0u;
            bs.pad1 = bs.pad2 = 0;
            cap = std::max(cap, stack_capacity(s));
        }
        BatchItem& bi = b->hItems[used++];
        bi.rays = it.rays;
        bi.out = it.out;
        bi.counts = it.counts;
        bi.scene = si;
        bi.mode = (it.flags & RT_ANY) ? kAny : kClosest;
        bi.start = 0;
        bi.n = it.n;
        total += it.n;
        if (total >= 0x80000000ull) return set_error(RT_ERR_ARG, "rt_batch_trace: at most 2^31-1 rays per batch");
    }
    if (used == 0) return RT_OK;
    const bool graph = (flags & RT_GRAPH) != 0;
    const bool sceneShapes = (flags & RT_SCENE_SHAPES) != 0;
    // Node format and launch shape of every scene.  Without RT_SCENE_SHAPES
    // the batch has one: the quantized records when every scene has them,
    // and the batch's launch shape.  With it, each scene its own: its node
    // format (rt_set_node_format, RT_EXACT_BOXES), its launch shape
    // (rt_set_launch) and its wave schedule.
    std::vector<uint64_t> sceneKey(nScenes);
    {
        bool allQ = !(flags & RT_EXACT_BOXES);
        for (uint32_t i = 0; i < nScenes; ++i) allQ = allQ && scenes[i]->qnodes;
        for (uint32_t i = 0; i < nScenes; ++i) {
            const rt_scene* sc = scenes[i];
            bool q;
            uint32_t bd, S, sched;
            if (sceneShapes) {
                q = use_quantized(sc, kClosest, flags) && !sc->exactBoxes;
                const bool qShape = q && !sc->launchSet;
                bd = qShape ? kQBlock : sc->blockSize;
                S = qShape ? kQWindow : sc->stackWindow;
                sched = sc->sched;
            } else {
                q = allQ;
                const bool qShape = q && !b->launchSet;
                bd = qShape ? kQBlock : b->blockSize;
                S = qShape ? kQWindow : b->stackWindow;
                sched = kDefaultSched;
            }
            if (q) b->hScenes[i].nodes = sc->qnodes;
            sceneKey[i] = (uint64_t)(q ? 1u : 0u) << 56 | (uint64_t)bd << 40 | (uint64_t)S << 32 | sched;
        }
    }
    // The items are ordered (stable) by segment, then by node memory, then
    // closest before any hit, into groups that the segment's kernel runs one
    // after another, each with the body specialised on its node memory and
    // mode.  A segment whose scenes share one node memory runs the kernel
    // specialised on it; otherwise the grouped kernel.  Item starts are the
    // running ray offsets within the segment.
    auto gkey = [&](const BatchItem& x) { return batch_group_key(b->hScenes[x.scene].M, x.mode); };
    std::stable_sort(b->hItems, b->hItems + used, [&](const BatchItem& x, const BatchItem& y) {
        const uint64_t kx = sceneKey[x.scene], ky = sceneKey[y.scene];
        return kx != ky ? kx < ky : gkey(x) < gkey(y);
    });
    std::vector<BatchSeg> segs;
    for (uint32_t i = 0; i < used; ++i) {
        BatchItem& bi = b->hItems[i];
        const uint64_t sk = sceneKey[bi.scene];
        if (segs.empty() || sceneKey[b->hItems[segs.back().i0].scene] != sk) {
            BatchSeg g;
            g.q = (sk >> 56) != 0;
            g.bd = (uint32_t)(sk >> 40) & 0xFFFFu;
            g.S = (uint32_t)(sk >> 32) & 0xFFu;
            g.sched = (uint32_t)sk;
            g.M = b->hScenes[bi.scene].M;
            g.i0 = i;
            segs.push_back(g);
        }
        BatchSeg& g = segs.back();
        g.i1 = i + 1;
        if (b->hScenes[bi.scene].M != g.M) g.M = kBatchMixed;
        g.maxM = std::max(g.maxM, b->hScenes[bi.scene].M);
        const uint32_t k = gkey(bi);
        if (g.nGroups == 0 || g.groupKey[g.nGroups - 1] != k) {
            if (g.nGroups == kMaxGroups)
                return set_error(RT_ERR_LAYOUT, "rt_batch_trace: more than 8 (node memory, mode) groups");
            g.groupKey[g.nGroups++] = k;
        }
        bi.start = (uint32_t)g.total;
        g.total += bi.n;
        g.groupEnd[g.nGroups - 1] = (uint32_t)g.total;
    }
    if (segs.size() > kMaxSegs)
        return set_error(RT_ERR_ARG, "rt_batch_trace: more than " + std::to_string(kMaxSegs) +
                                         " distinct scene launch shapes");
    // mixed node memories all up to 8: the grouped kernel without the larger bodies
    for (BatchSeg& g : segs)
        if (g.M == kBatchMixed && g.maxM <= 8) g.M = kBatchMixedSmall;
    // A captured graph with the same segment shapes is re-armed; it keeps the
    // larger LDS table reservation and spill capacity it was made with.
    const bool reuse = graph && b->exec && b->graphStream == st && b->graphSegs.size() == segs.size();
    size_t spillWords = 0;
    uint32_t maxStack = 0;
    for (size_t i = 0; i < segs.size(); ++i) {
        BatchSeg& g = segs[i];
        const BatchSeg* old = reuse ? &b->graphSegs[i] : nullptr;
        if (old && (old->q != g.q || old->bd != g.bd || old->S != g.S || old->M != g.M)) old = nullptr;
        // the scene and item tables go to LDS
        g.table = batch_table_bytes(nScenes, g.i1 - g.i0);
        if (old && old->table >= g.table) g.table = old->table;
        const int occ = trace_occupancy(g.M, kBatch, g.S, g.bd, g.q, g.table);
        if (occ <= 0)
            return set_error(RT_ERR_ARG, "rt_batch_trace: the tables of " + std::to_string(nScenes) + " scenes and " +
                                             std::to_string(g.i1 - g.i0) + " items (" + std::to_string(g.table) +
                                             " bytes) do not fit the kernel's LDS beside a " + std::to_string(g.S) +
                                             "-entry stack window of " + std::to_string(g.bd) + " threads");
        const uint64_t full = (uint64_t)b->cus * (uint64_t)occ;
        g.grid = graph ? (uint32_t)full
                       : (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(full, (g.total + g.bd - 1) / g.bd));
        g.spillCap = cap > g.S ? cap - g.S : 0;
        if (old && old->spillCap > g.spillCap && old->grid == g.grid) g.spillCap = old->spillCap;
        // the segments run at once: each has a spill range of its own
        g.spillOff = spillWords;
        spillWords += (size_t)g.grid * g.bd * g.spillCap * (g.q ? 2 : 1);
        maxStack = std::max(maxStack, g.spillCap + g.S);
    }
    int rc;
    if ((rc = ensure(reinterpret_cast<void**>(&b->spill), b->spillBytes, std::max<size_t>(64, spillWords * 4),
                     "hipMalloc(spill)")) != RT_OK)
        return rc;
    HIP_TRY(hipMemcpyAsync(b->dItems, b->hItems, sizeof(BatchItem) * used, hipMemcpyHostToDevice, st), "items");
    HIP_TRY(hipMemcpyAsync(b->dScenes, b->hScenes, sizeof(BatchScene) * nScenes, hipMemcpyHostToDevice, st), "scenes");
    std::vector<KParams> ps(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) {
        const BatchSeg& g = segs[i];
        KParams& p = ps[i];
        p = KParams{};
        p.n = (uint32_t)g.total;
        p.spill = b->spill + g.spillOff;
        p.spillCap = g.spillCap;
        p.work = b->work + (size_t)kWorkWords * i;
        p.items = b->dItems + g.i0;
        p.scenes = b->dScenes;
        p.nItems = g.i1 - g.i0;
        p.nScenes = nScenes;
        p.nGroups = g.nGroups;
        for (uint32_t k = 0; k < g.nGroups; ++k) {
            p.groupKey[k] = g.groupKey[k];
            p.groupEnd[k] = g.groupEnd[k];
        }
        p.chunk = ray_chunk(g.total, g.grid, g.bd);
        p.sched = g.sched;
    }
    b->maxStack = maxStack;
    if (graph) {
        bool same = reuse && b->graphSpill == b->spill;
        for (size_t i = 0; same && i < segs.size(); ++i) same = segs[i].same_launch(b->graphSegs[i]);
        if (!same) {
            if ((rc = build_graph(b, st, segs, ps)) != RT_OK) return rc;
        } else {
            // re-arm: same graph, new kernel arguments (ray counts, item tables)
            for (size_t i = 0; i < segs.size(); ++i) {
                hipKernelNodeParams kp = b->traceNodeParams[i];
                KParams copy = ps[i];
                void* args[] = {&copy};
                kp.kernelParams = args;
                kp.extra = nullptr;
                HIP_TRY(hipGraphExecKernelNodeSetParams(b->exec, b->traceNodes[i], &kp),
                        "hipGraphExecKernelNodeSetParams");
            }
        }
        HIP_TRY(hipEventRecord(b->ev0, st), "hipEventRecord");
        HIP_TRY(hipGraphLaunch(b->exec, st), "hipGraphLaunch");
        HIP_TRY(hipEventRecord(b->ev1, st), "hipEventRecord");
        ++b->graphLaunches;
    } else {
        HIP_TRY(hipEventRecord(b->ev0, st), "hipEventRecord");
        if ((rc = launch_segments(b, st, segs, ps)) != RT_OK) return rc;
        HIP_TRY(hipEventRecord(b->ev1, st), "hipEventRecord");
    }
    b->pending = true;
    b->raysTraced += total;
    if ((flags & RT_NO_SYNC) && joinNull) {
        const int jr = null_stream_after(st, b->joinEv);
        if (jr != RT_OK) return jr;
    }
    if (!(flags & RT_NO_SYNC)) {
        HIP_TRY(hipEventSynchronize(b->ev1), "batched trace kernel");
        b->pending = false;
        float ms = 0;
        if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->lastMs = ms;
    }
    return RT_OK;
}

extern "C" int rt_batch_stats(rt_batch_t b, rt_stats_t* o) {
    if (!b || !o) return set_error(RT_ERR_ARG, "rt_batch_stats: null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    if (b->pending && hipEventQuery(b->ev1) == hipSuccess) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->lastMs = ms;
    }
    std::memset(o, 0, sizeof(*o));
    o->bytes_resident = b->spillBytes + b->maxItems * (sizeof(BatchItem) + sizeof(BatchScene));
    o->last_kernel_ms = b->lastMs;
    o->rays_traced = b->raysTraced;
    o->block_size = b->blockSize;
    o->stack_window = b->stackWindow;
    o->max_stack = b->maxStack;
    // the guard counter accumulates on the device over the batch's launches
    if (b->pending) HIP_TRY(hipEventSynchronize(b->ev1), "hipEventSynchronize(stats)");
    std::vector<uint32_t> words((size_t)kWorkWords * kMaxSegs);
    HIP_TRY(hipMemcpy(words.data(), b->work, words.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy(errors)");
    uint64_t errs = 0;
    for (uint32_t i = 0; i < kMaxSegs; ++i) errs += words[(size_t)kWorkWords * i + kErrWord];
    o->error_rays = (uint32_t)std::min<uint64_t>(errs, 0xFFFFFFFFu);
    return RT_OK;
}

extern "C" int rt_batch_release(rt_batch_t b) {
    if (!b) return RT_OK;
    if (b->pending) (void)hipEventSynchronize(b->ev1);
    free_batch(b);
    return RT_OK;
}
