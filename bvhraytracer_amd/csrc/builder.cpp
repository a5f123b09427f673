// N-ary SAH BVH builder and FastNode flattener (host side of the path).
//
// Restates, step for step, the reference builder:
//   Bvh::Bvh / Bvh::recursiveBvh        src/accelerationStructure/bvh.cpp:28-40
//   chooseAxisAndSort                   src/accelerationStructure/aabb.h:40-69
//   Aabb::calculateBounds / getCenter   aabb.h:113-146
//   sweepRight / sweepLeft              aabb.h:205-222 (left sweep off by one)
//   Node::sah (step-function leaf cost) src/accelerationStructure/node.h:66-100
//   computerBestSplit(Sort)             src/accelerationStructure/aabb.cpp:4-69
//   Aabb::recursiveBvh                  aabb.cpp:71-206
//   calculateTraverseOrderEachAxis      src/accelerationStructure/node.cpp:254-275
//   Bvh::bvhAnalysis (depth, counts)    bvh.cpp:338-389
//   FastNodeManager ctor + customTreeOrder  fastNodeManager.h:1495-1688
// and RayTracer::preparePrimitives (rayTracer.cpp:747-754).  All float
// arithmetic is FP32 in the reference's evaluation order (build with
// -ffp-contract=off).  The reference sorts with an unstable parallel
// std::sort; this builder uses std::stable_sort so the topology is
// deterministic (equal keys keep their current relative order).
#include <algorithm>
#include <atomic>
#include <limits>
#include <memory>
#include <thread>

#include "host_common.h"

namespace rt {

namespace {

struct Prim {
    float bmin[3], bmax[3], c[3];
    uint32_t id;
};

// glm::min / glm::max component semantics: min(x,y) = (y < x) ? y : x,
// max(x,y) = (x < y) ? y : x.
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }

struct Box {
    float mn[3] = {222222.0f, 222222.0f, 222222.0f};
    float mx[3] = {-222222.0f, -222222.0f, -222222.0f};
    void add(const Prim& p) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = gmin(mn[a], p.bmin[a]);
            mx[a] = gmax(mx[a], p.bmax[a]);
        }
    }
    // union with a box built by add() from the same initial corners: gmin /
    // gmax of two non-NaN values (add() never takes a NaN in), and the
    // earlier box wins ties, so folding chunk boxes in order equals adding
    // the primitives one by one
    void merge(const Box& o) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = gmin(mn[a], o.mn[a]);
            mx[a] = gmax(mx[a], o.mx[a]);
        }
    }
    // Aabb::getSurfaceArea (aabb.h:130-134)
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    }
};

// maxDimension (glmUtil.h:7-11): ties go to the later axis.
inline int max_dimension(float x, float y, float z) {
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

// Node::sah(invArea, leafTarget) for primCount != 0 (node.h:78-91)
inline float sah_leaf(size_t count, float area, float invArea, int leafTarget) {
    size_t steps = ((count - 1) / (size_t)leafTarget) + 1;
    return (float)steps * area * invArea * (float)leafTarget;
}

// Ranges at least this long are swept, scanned and sorted on several threads
// (only near the root: below it the subtrees already run in parallel).
constexpr size_t kParallelRange = 1u << 17;

// fn(k) for k in [0, parts): parts - 1 new threads, the caller runs part 0.
template <class F>
void run_parts(size_t parts, F fn) {
    std::vector<std::thread> ts;
    for (size_t k = 1; k < parts; ++k) ts.emplace_back([&fn, k] { fn(k); });
    fn(0);
    for (auto& t : ts) t.join();
}

// Stable sort of [first, last) on up to `threads` threads: chunks are
// stable-sorted in parallel, then adjacent runs are merged pairwise with
// std::merge (which takes equal keys from the left run first).  A stable sort
// has exactly one result for a given order and comparator, so this equals
// std::stable_sort bit for bit.
template <class It, class Less>
void parallel_stable_sort(It first, It last, Less less, int threads) {
    using T = typename std::iterator_traits<It>::value_type;
    const size_t n = (size_t)(last - first);
    const size_t parts = std::min<size_t>((size_t)std::max(1, threads), 16);
    if (parts < 2 || n < kParallelRange) {
        std::stable_sort(first, last, less);
        return;
    }
    std::vector<size_t> cut(parts + 1);
    for (size_t i = 0; i <= parts; ++i) cut[i] = n * i / parts;
    {
        std::vector<std::thread> ts;
        for (size_t i = 0; i < parts; ++i)
            ts.emplace_back([&, i] { std::stable_sort(first + cut[i], first + cut[i + 1], less); });
        for (auto& t : ts) t.join();
    }
    std::vector<T> tmp(n);
    T* src = &*first;
    T* dst = tmp.data();
    bool inTmp = false;
    for (size_t width = 1; width < parts; width *= 2) {
        std::vector<std::thread> ts;
        for (size_t i = 0; i < parts; i += 2 * width) {
            const size_t a = cut[i], m = cut[std::min(i + width, parts)], b = cut[std::min(i + 2 * width, parts)];
            ts.emplace_back([=] { std::merge(src + a, src + m, src + m, src + b, dst + a, less); });
        }
        for (auto& t : ts) t.join();
        std::swap(src, dst);
        inTmp = !inTmp;
    }
    if (inTmp) std::copy(tmp.begin(), tmp.end(), first);
}

struct BNode {
    Box box;
    uint32_t begin = 0, end = 0;  // primitive range (empty for inner nodes)
    uint8_t depth = 0;
    std::vector<std::unique_ptr<BNode>> children;
    std::vector<int8_t> order[3];
    mutable uint32_t fid = 0;  // position in the flattened node array
};

struct Builder {
    std::vector<Prim>& prims;
    unsigned N, L;
    bool sortEachSplit;
    int leafSplitOption;
    float sahFactor;
    std::atomic<int> spareThreads;
    int threads;

    Builder(std::vector<Prim>& p, unsigned n, unsigned l, bool s, int o, float f, int threads)
        : prims(p), N(n), L(l), sortEachSplit(s), leafSplitOption(o), sahFactor(f),
          spareThreads(threads - 1), threads(threads) {}

    size_t parts_for(size_t n) const {
        return (n >= kParallelRange && threads > 1) ? (size_t)std::min(threads, 16) : 1;
    }

    // chooseAxisAndSort (aabb.h:40-69).  The centroid bounds and the
    // is-sorted test are exact in chunks (see Box::merge; sortedness is
    // per chunk plus the chunk seams).
    int choose_axis_and_sort(uint32_t b, uint32_t e) {
        const size_t n = e - b, parts = parts_for(n);
        std::vector<Box> cb(parts);
        run_parts(parts, [&](size_t k) {
            Box& x = cb[k];
            for (size_t i = b + n * k / parts; i < b + n * (k + 1) / parts; ++i)
                for (int a = 0; a < 3; ++a) {
                    x.mn[a] = gmin(x.mn[a], prims[i].c[a]);
                    x.mx[a] = gmax(x.mx[a], prims[i].c[a]);
                }
        });
        Box c;
        for (const Box& x : cb) c.merge(x);
        int axis = max_dimension(c.mx[0] - c.mn[0], c.mx[1] - c.mn[1], c.mx[2] - c.mn[2]);
        auto less = [axis](const Prim& p1, const Prim& p2) { return p1.c[axis] < p2.c[axis]; };
        std::vector<char> ok(parts, 1);
        run_parts(parts, [&](size_t k) {
            const size_t lo = b + n * k / parts, hi = b + n * (k + 1) / parts;
            ok[k] = std::is_sorted(prims.begin() + lo, prims.begin() + hi, less) &&
                    (hi == e || !less(prims[hi], prims[hi - 1]));
        });
        if (std::find(ok.begin(), ok.end(), 0) != ok.end()) {
            // large ranges are only sorted near the root, before the
            // subtrees take the threads
            if (parts > 1) parallel_stable_sort(prims.begin() + b, prims.begin() + e, less, threads);
            else std::stable_sort(prims.begin() + b, prims.begin() + e, less);
        }
        return axis;
    }

    // computerBestSplit(Sort) (aabb.cpp:4-69).  Returns the split position.
    // Long ranges run the two sweeps in chunks: the box entering a chunk is
    // the fold of the chunks before it (after it, for the right sweep), which
    // equals the running box of the one-by-one sweep (Box::merge), so every
    // metric[m] is bit-identical; the first minimum is taken per chunk and
    // then across chunks in order.
    uint32_t best_split(uint32_t b, uint32_t e, float invArea, float& cost) {
        if (sortEachSplit) choose_axis_and_sort(b, e);
        const size_t size = e - b, cnt = size - 1;
        std::vector<float> metric(cnt, 0.0f);
        const size_t parts = parts_for(size);
        auto lo = [&](size_t k) { return cnt * k / parts; };
        std::vector<Box> lstart(parts), rstart(parts);
        if (parts > 1) {
            std::vector<Box> chunk(parts);
            run_parts(parts, [&](size_t k) {
                for (size_t m = lo(k); m < lo(k + 1); ++m) chunk[k].add(prims[b + m + 1]);
            });
            Box acc;
            for (size_t k = 0; k < parts; ++k) {
                lstart[k] = acc;
                acc.merge(chunk[k]);
            }
            acc = Box();
            for (size_t k = parts; k-- > 0;) {
                rstart[k] = acc;
                acc.merge(chunk[k]);
            }
        }
        std::vector<size_t> best(parts, 0);
        run_parts(parts, [&](size_t k) {
            Box left = lstart[k];  // sweepRight reads *(++primitiveEnd): prims b+1 .. b+m+1
            for (size_t m = lo(k); m < lo(k + 1); ++m) {
                left.add(prims[b + m + 1]);
                metric[m] += sah_leaf(m + 1, left.area(), invArea, (int)L);
            }
            Box right = rstart[k];  // sweepLeft reads *(--primitiveBegin): prims b+m+1 .. e-1
            for (size_t m = lo(k + 1); m-- > lo(k);) {
                right.add(prims[b + m + 1]);
                metric[m] += sah_leaf(size - 1 - m, right.area(), invArea, (int)L);
            }
            size_t bm = lo(k);  // std::min_element: first minimum
            for (size_t m = lo(k) + 1; m < lo(k + 1); ++m)
                if (metric[m] < metric[bm]) bm = m;
            best[k] = bm;
        });
        size_t bi = best[0];
        for (size_t k = 1; k < parts; ++k)
            if (lo(k) < lo(k + 1) && metric[best[k]] < metric[bi]) bi = best[k];
        cost = metric[bi];
        return b + (uint32_t)bi + 1;
    }

    void bounds_of(BNode& n) {
        n.box = Box();
        for (uint32_t i = n.begin; i < n.end; ++i) n.box.add(prims[i]);
    }

    // Aabb::recursiveBvh (aabb.cpp:71-206) + Node::recursiveBvh (node.cpp:23-30)
    void recurse(BNode& node) {
        const size_t count = node.end - node.begin;
        if (count <= L && (leafSplitOption == 0 || leafSplitOption == 2)) return;
        choose_axis_and_sort(node.begin, node.end);
        struct Iv {
            uint32_t b, e;
        };
        std::vector<Iv> work{{node.begin, node.end}};
        size_t bestI = 0;
        const float invArea = 1.0f / node.box.area();
        for (unsigned br = 0; br + 1 < N; ++br) {
            size_t primCounter = 0;
            for (size_t i = 0; i < work.size(); ++i)
                if (primCounter < work[i].e - work[i].b) {
                    primCounter = work[i].e - work[i].b;
                    bestI = i;
                }
            if (primCounter <= L) break;
            float cost;
            uint32_t split = best_split(work[bestI].b, work[bestI].e, invArea, cost);
            Iv p2{split, work[bestI].e};
            work[bestI].e = split;
            work.insert(work.begin() + bestI + 1, p2);
        }
        if (leafSplitOption == 1 || leafSplitOption == 2) {
            for (size_t br = work.size() - 1; br + 1 < N; ++br) {
                float bestImp = 0.0f;
                long bestId = -1;
                uint32_t bestSplit = 0;
                for (size_t i = 0; i < work.size(); ++i) {
                    int cnt = (int)(work[i].e - work[i].b);
                    if (cnt == 1) continue;
                    float leafCost = (float)cnt;
                    float splitCost = 0.0f;
                    uint32_t split = best_split(work[i].b, work[i].e, invArea, splitCost);
                    splitCost += sahFactor;
                    if (leafCost - splitCost > bestImp) {
                        bestImp = leafCost - splitCost;
                        bestId = (long)i;
                        bestSplit = split;
                    }
                }
                if (bestImp == 0.0f) break;
                Iv p2{bestSplit, work[bestId].e};
                work[bestId].e = bestSplit;
                work.insert(work.begin() + bestId + 1, p2);
            }
        }
        if (work.size() == 1) return;
        for (auto& iv : work) {
            auto c = std::make_unique<BNode>();
            c->begin = iv.b;
            c->end = iv.e;
            c->depth = (uint8_t)(node.depth + 1);
            bounds_of(*c);
            node.children.push_back(std::move(c));
        }
        if (sortEachSplit) traverse_order(node);
        node.begin = node.end;  // primitiveBegin = primitiveEnd: inner node
        // children are independent: build the big ones on spare threads
        std::vector<std::thread> ts;
        for (auto& c : node.children) {
            BNode* cp = c.get();
            if (cp->end - cp->begin > 20000 && spareThreads.fetch_sub(1) > 0) {
                ts.emplace_back([this, cp] {
                    recurse(*cp);
                    spareThreads.fetch_add(1);
                });
            } else {
                if (cp->end - cp->begin > 20000) spareThreads.fetch_add(1);
                recurse(*cp);
            }
        }
        for (auto& t : ts) t.join();
    }

    // Node::calculateTraverseOrderEachAxis (node.cpp:254-275); the key is
    // Aabb::getCenter = (max + min) + 0.5f (aabb.h:142-146), ties by index.
    void traverse_order(BNode& node) {
        const size_t cc = node.children.size();
        for (int face = 0; face < 3; ++face) {
            std::vector<std::pair<float, int>> t(cc);
            for (size_t i = 0; i < cc; ++i) {
                const Box& bx = node.children[i]->box;
                t[i] = {(bx.mx[face] + bx.mn[face]) + 0.5f, (int)i};
            }
            std::sort(t.begin(), t.end());
            node.order[face].clear();
            for (auto& p : t) node.order[face].push_back((int8_t)p.second);
        }
    }
};

void flatten(const BNode& root, const std::vector<Prim>& prims, const float* tris, unsigned G, int threads,
             HostBvh& out) {
    // customTreeOrder (fastNodeManager.h:1675-1688): root, then every node's
    // children as one block, recursively in child order.  Entering a node
    // appends its children; the iteration below is that recursion unrolled.
    std::vector<const BNode*> nodes{&root};
    struct Frame {
        const BNode* n;
        size_t next;
    };
    root.fid = 0;
    auto append_children = [&nodes](const BNode* n) {
        for (auto& c : n->children) {
            c->fid = (uint32_t)nodes.size();
            nodes.push_back(c.get());
        }
    };
    append_children(&root);
    std::vector<Frame> fs{{&root, 0}};
    while (!fs.empty()) {
        Frame& f = fs.back();
        if (f.next < f.n->children.size()) {
            const BNode* c = f.n->children[f.next++].get();
            append_children(c);
            fs.push_back({c, 0});
        } else {
            fs.pop_back();
        }
    }
    const uint32_t M = out.M, Lm = out.Lm;
    const size_t nn = nodes.size();
    auto id_of = [](const BNode* p) { return p->fid; };
    // leaf blocks are appended in node order: leaf ordinal of each node
    std::vector<uint32_t> leafOrd(nn, 0);
    uint32_t leaves = 0;
    for (size_t i = 0; i < nn; ++i)
        if (nodes[i]->end > nodes[i]->begin) leafOrd[i] = leaves++;
    out.bounds.assign(nn * 6 * M, 0.0f);
    out.base.assign(nn, 0);
    out.order.assign(nn * 3 * M, 0);
    out.childType.assign(nn, 0);
    out.tri.assign((size_t)leaves * 9 * Lm, 0.0f);
    out.triIds.assign((size_t)leaves * Lm, -1);
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    // nodes are written independently: ranges of the node array on threads
    const size_t parts = (nn >= (1u << 15) && threads > 1) ? (size_t)std::min(threads, 16) : 1;
    std::vector<double> depthSums(parts, 0.0);
    std::vector<uint32_t> triCounts(parts, 0), maxDepths(parts, 0);
    run_parts(parts, [&](size_t k) {
        for (size_t i = nn * k / parts; i < nn * (k + 1) / parts; ++i) {
            const BNode& n = *nodes[i];
            float* bd = &out.bounds[i * 6 * M];
            const size_t cc = n.children.size();
            if (cc > 0) {
                out.base[i] = id_of(n.children[0].get());
                const uint32_t nrs = (uint32_t)(((cc - 1) / G + 1) * G);
                for (int a = 0; a < 3; ++a)
                    for (size_t j = 0; j < cc; ++j) {
                        bd[a * nrs + j] = n.children[j]->box.mn[a];
                        bd[nrs * 3 + a * nrs + j] = n.children[j]->box.mx[a];
                    }
                if (nrs != M) {
                    bd[6 * M - 2] = (float)nrs;
                    bd[6 * M - 1] = qnan;
                }
                uint16_t ct = 0;
                for (size_t j = 0; j < cc; ++j)
                    if (!n.children[j]->children.empty()) ct |= (uint16_t)(1u << j);
                out.childType[i] = ct;
                for (int a = 0; a < 3; ++a)
                    for (uint32_t p = 0; p < M; ++p)
                        out.order[i * 3 * M + a * M + p] = (p < n.order[0].size()) ? n.order[a][p] : (int8_t)0;
            }
            const uint32_t pc = n.end - n.begin;
            if (pc > 0) {
                const size_t pBegin = (size_t)leafOrd[i] * 9 * Lm;
                const uint32_t lrs = ((pc - 1) / G + 1) * G;
                for (int v = 0; v < 3; ++v)
                    for (int a = 0; a < 3; ++a)
                        for (uint32_t j = 0; j < pc; ++j)
                            out.tri[pBegin + (v * 3 + a) * lrs + j] = tris[(size_t)prims[n.begin + j].id * 9 + v * 3 + a];
                if (lrs != Lm) {
                    out.tri[pBegin + Lm * 9 - 2] = (float)lrs;
                    out.tri[pBegin + Lm * 9 - 1] = qnan;
                }
                bd[0] = qnan;
                out.base[i] = (uint32_t)pBegin;
                for (uint32_t j = 0; j < pc; ++j)
                    out.triIds[(size_t)leafOrd[i] * Lm + j] = (int32_t)prims[n.begin + j].id;
                depthSums[k] += n.depth;
                triCounts[k] += pc;
                maxDepths[k] = std::max<uint32_t>(maxDepths[k], n.depth);
            }
        }
    });
    double depthSum = 0;  // integer depths: exact in any order
    uint32_t triCount = 0, maxDepth = 0;
    for (size_t k = 0; k < parts; ++k) {
        depthSum += depthSums[k];
        triCount += triCounts[k];
        maxDepth = std::max(maxDepth, maxDepths[k]);
    }
    out.averageDepth = leaves ? (float)(depthSum / leaves) : 0.0f;
    out.triangleCount = triCount;
    out.maxDepth = maxDepth;
}

}  // namespace

int build_bvh(const float* tris, uint32_t n, const rt_build_params& p, HostBvh& out) {
    if (!tris || n == 0) return set_error(RT_ERR_ARG, "rt_build: empty triangle list");
    if (p.gang != 4 && p.gang != 8 && p.gang != 16)
        return set_error(RT_ERR_UNSUPPORTED, "rt_build: gang size must be 4, 8 or 16");
    if (p.branching < 2 || p.leaf_size < 1)
        return set_error(RT_ERR_ARG, "rt_build: branching factor >= 2 and leaf size >= 1 required");
    const uint32_t G = p.gang;
    out.N = p.branching;
    out.L = p.leaf_size;
    out.G = G;
    out.M = (p.branching + G - 1) / G * G;
    out.Lm = (p.leaf_size + G - 1) / G * G;
    // the node / leaf memory tests (rayTracer.cpp:41-52,630-640) flatten into
    // larger records: a multiple of G at least the minimum
    if (p.node_memory) {
        if (p.node_memory < out.M || p.node_memory % G)
            return set_error(RT_ERR_ARG, "rt_build: node_memory must be a multiple of the gang size >= roundup(N, G)");
        out.M = p.node_memory;
    }
    if (p.leaf_memory) {
        if (p.leaf_memory < out.Lm || p.leaf_memory % G)
            return set_error(RT_ERR_ARG, "rt_build: leaf_memory must be a multiple of the gang size >= roundup(L, G)");
        out.Lm = p.leaf_memory;
    }
    if (out.M > 16 || out.Lm > 16)
        return set_error(RT_ERR_UNSUPPORTED, "rt_build: node and leaf memory are limited to 16");
    std::vector<Prim> prims(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = tris + (size_t)i * 9;
        Prim& q = prims[i];
        // Triangle::updateBounds (triangle.cpp:41-49)
        for (int a = 0; a < 3; ++a) {
            float lo = gmin(t[a], t[3 + a]), hi = gmax(t[a], t[3 + a]);
            q.bmin[a] = gmin(lo, t[6 + a]);
            q.bmax[a] = gmax(hi, t[6 + a]);
            q.c[a] = q.bmax[a] * 0.5f + q.bmin[a] * 0.5f;
        }
        q.id = i;
    }
    int threads = p.threads ? (int)p.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    Builder b(prims, p.branching, p.leaf_size, p.sort_each_split != 0, p.leaf_split_option,
              p.sah_factor, threads);
    b.choose_axis_and_sort(0, n);  // RayTracer::preparePrimitives
    BNode root;
    root.begin = 0;
    root.end = n;
    b.bounds_of(root);
    b.recurse(root);
    flatten(root, prims, tris, G, threads, out);
    return RT_OK;
}

}  // namespace rt
