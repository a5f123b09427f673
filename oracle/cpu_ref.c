/*
 * ORACLE — test infrastructure, not product code (see cpu_ref.h).
 * Build: gcc -O3 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
 * Every float operation below is FP32 in the order the reference states it;
 * the ISPC sse4-i32x4 target has no FMA, so contraction is disabled.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_STACK_CAP 1024

/* The node and leaf tests the traversals call.  cpu_simd.c includes this
 * file with its 4-wide forms of the three tests defined in their place. */
#ifndef ORC_NODE_TEST
#define ORC_NODE_TEST orc_aabb_intersect
#define ORC_LEAF_TEST orc_tri_intersect
#define ORC_LEAF_ANY orc_tri_anyhit
#endif

/* ISPC max()/min() on sse4 lower to maxps/minps: (a > b) ? a : b and
 * (a < b) ? a : b — the second operand wins on NaN and on equal values. */
static inline float smax(float a, float b) { return (a > b) ? a : b; }
static inline float smin(float a, float b) { return (a < b) ? a : b; }

/* aabbIntersect{k} (rayTracerTemplate.ispc:170-234).  Writes dist[j] = t0
 * for hit slots only; misses leave dist[j] untouched. */
int orc_aabb_intersect(const float* bounds, int k, float* dist, const float* ri) {
    int result = 0;
    for (int index = 0; index < k; ++index) {
        const float bminx = bounds[index], bminy = bounds[index + k], bminz = bounds[index + k * 2];
        const float bmaxx = bounds[index + k * 3], bmaxy = bounds[index + k * 4], bmaxz = bounds[index + k * 5];
        float t0 = 0.0f, t1 = ri[0];
        float tnx = (bminx - ri[1]) * ri[4], tny = (bminy - ri[2]) * ri[5], tnz = (bminz - ri[3]) * ri[6];
        float tfx = (bmaxx - ri[1]) * ri[4], tfy = (bmaxy - ri[2]) * ri[5], tfz = (bmaxz - ri[3]) * ri[6];
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
        if (t0 <= t1) {
            result = 1;
            dist[index] = t0;
        }
    }
    return result;
}

/* One Moeller-Trumbore test exactly as rayTracerTemplate.ispc:49-75.
 * Returns the `hit` flag (a NaN b1/b2/t does not clear it). */
static inline int mt32(const float p0[3], const float p1[3], const float p2[3], const float* ri,
                       float* tout, float* b1out, float* b2out) {
    const float ox = ri[1], oy = ri[2], oz = ri[3];
    const float dx = ri[7], dy = ri[8], dz = ri[9];
    const float e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
    const float e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
    /* s1 = cross(dir, e2) */
    const float s1x = dy * e2z - dz * e2y;
    const float s1y = dz * e2x - dx * e2z;
    const float s1z = dx * e2y - dy * e2x;
    const float divisor = s1x * e1x + s1y * e1y + s1z * e1z;
    int hit = 1;
    if (divisor == 0.0f) hit = 0;
    const float inv = 1.0f / divisor;
    const float ddx = ox - p0[0], ddy = oy - p0[1], ddz = oz - p0[2];
    const float b1 = (ddx * s1x + ddy * s1y + ddz * s1z) * inv;
    if (b1 < 0.0f || b1 > 1.0f) hit = 0;
    /* s2 = cross(d, e1) */
    const float s2x = ddy * e1z - ddz * e1y;
    const float s2y = ddz * e1x - ddx * e1z;
    const float s2z = ddx * e1y - ddy * e1x;
    const float b2 = (dx * s2x + dy * s2y + dz * s2z) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) hit = 0;
    const float t = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
    if (t < 0.0f || t > ri[0]) hit = 0;
    *tout = t;
    *b1out = b1;
    *b2out = b2;
    return hit;
}

static inline void load_tri(const float* tri, uint32_t tri_id, int k, int index, float p0[3],
                            float p1[3], float p2[3]) {
    const float* b = tri + tri_id + index;
    p0[0] = b[0];
    p0[1] = b[k];
    p0[2] = b[k * 2];
    p1[0] = b[k * 3];
    p1[1] = b[k * 4];
    p1[2] = b[k * 5];
    p2[0] = b[k * 6];
    p2[1] = b[k * 7];
    p2[2] = b[k * 8];
}

/* triIntersect{k} (rayTracerTemplate.ispc:25-108) with gang size G: slot
 * `index` runs on program instance index % G; each instance keeps its first
 * strict minimum (tStore starts at 1e6); bestT = reduce_min; the result is
 * reduce_max of origIndex over instances whose tStore == bestT. */
int orc_tri_intersect(const float* tri, uint32_t tri_id, int k, float* ri, int G) {
    float tStore[16];
    int origIndex[16];
    int result = 0;
    for (int g = 0; g < G; ++g) {
        tStore[g] = 1000000.0f;
        origIndex[g] = -1;
    }
    for (int index = 0; index < k; ++index) {
        float p0[3], p1[3], p2[3], t, b1, b2;
        load_tri(tri, tri_id, k, index, p0, p1, p2);
        if (mt32(p0, p1, p2, ri, &t, &b1, &b2)) {
            const int lane = index % G;
            if (tStore[lane] > t) {
                result = 1;
                tStore[lane] = t;
                origIndex[lane] = index;
            }
        }
    }
    if (!result) return -1;
    float bestT = tStore[0];
    for (int g = 1; g < G; ++g) bestT = (tStore[g] < bestT) ? tStore[g] : bestT;
    ri[0] = bestT;
    int idx = -1;
    for (int g = 0; g < G; ++g)
        if (tStore[g] == bestT && origIndex[g] > idx) idx = origIndex[g];
    return idx;
}

/* triAnyHit{k} (rayTracerTemplate.ispc:111-168) */
int orc_tri_anyhit(const float* tri, uint32_t tri_id, int k, const float* ri) {
    int result = 0;
    for (int index = 0; index < k; ++index) {
        float p0[3], p1[3], p2[3], t, b1, b2;
        load_tri(tri, tri_id, k, index, p0, p1, p2);
        if (mt32(p0, p1, p2, ri, &t, &b1, &b2)) result = 1;
    }
    return result;
}

/* callIspcTemplateNotConst (fastNodeManager.h:271-275): 4, 8, 12 pick their
 * variant; every other loop count runs the 16 variant. */
int orc_dispatch_k(int k) { return (k == 4 || k == 8 || k == 12) ? k : 16; }

/* getNodeLoopCount (fastNodeManager.h:1748-1763) */
int orc_node_loop_count(const orc_bvh* b, uint32_t node) {
    const uint32_t M = b->M;
    if (M == b->G) return orc_dispatch_k((int)M);
    const float* bd = b->bounds + (size_t)node * 6 * M;
    int k = (int)M;
    if (isnan(bd[M * 6 - 1])) k = (int)bd[M * 6 - 2];
    return orc_dispatch_k(k);
}

/* getLeafLoopCount (fastNodeManager.h:1728-1746) */
int orc_leaf_loop_count(const orc_bvh* b, uint32_t pib) {
    const uint32_t Lm = b->Lm;
    if (Lm == b->G) return orc_dispatch_k((int)Lm);
    int k = (int)Lm;
    const uint64_t end = (uint64_t)pib + Lm * 9;
    if (isnan(b->tri[end - 1])) k = (int)b->tri[end - 2];
    return orc_dispatch_k(k);
}

/* maxAbsDimension (glmUtil.h:18-23) */
static inline int max_abs_dim(float x, float y, float z) {
    x = fabsf(x);
    y = fabsf(y);
    z = fabsf(z);
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

static inline int is_leaf(const orc_bvh* b, uint32_t id) {
    return isnan(b->bounds[(size_t)id * 6 * b->M]);
}

/* Visit recorder of orc_visit_sequences: FastNode ids in visit order,
 * leaves with bit 31 set; len counts every visit, ids keeps the first cap. */
typedef struct orc_seq {
    uint32_t* ids;
    uint32_t cap, len;
} orc_seq;

static inline void count_add(orc_counts* c, orc_seq* q, uint32_t id, int leaf, int k) {
    if (q) {
        if (q->len < q->cap) q->ids[q->len] = id | (leaf ? 0x80000000u : 0u);
        q->len++;
    }
    if (!c) return;
    if (leaf) {
        c->leaves++;
        c->leaf_slots += (uint32_t)k;
    } else {
        c->nodes++;
        c->node_slots += (uint32_t)k;
    }
}

/* FastNodeManager::intersect (fastNodeManager.h:1328-1418) */
static int intersect_impl(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index, int32_t* tri_index,
                          orc_counts* counts, uint32_t* diag, orc_seq* q) {
    uint32_t stack[ORC_STACK_CAP];
    int sp = 1, result = 0;
    uint32_t dg = 0;
    float dist[16];
    float* ri = (float*)ray;
    const uint32_t M = b->M;
    stack[0] = 0;
    for (uint32_t i = 0; i < M; ++i) dist[i] = NAN;
    const int code = max_abs_dim(ray->dx, ray->dy, ray->dz);
    const float dcode = code == 0 ? ray->dx : (code == 1 ? ray->dy : ray->dz);
    const int reverse = dcode <= 0.0f;
    while (sp != 0) {
        const uint32_t id = stack[--sp];
        if (id >= b->node_count) {
            dg |= ORC_DIAG_BOGUS;
            continue;
        }
        if (is_leaf(b, id)) {
            const int k = orc_leaf_loop_count(b, b->base[id]);
            count_add(counts, q, id, 1, k);
            const int r = ORC_LEAF_TEST(b->tri, b->base[id], k, ri, (int)b->G);
            if (r != -1) {
                result = 1;
                *leaf_index = b->base[id];
                *tri_index = r;
            }
        } else {
            const int k = orc_node_loop_count(b, id);
            count_add(counts, q, id, 0, k);
            const float* bd = b->bounds + (size_t)id * 6 * M;
            if (ORC_NODE_TEST(bd, k, dist, ri)) {
                const int8_t* ord = b->order + (size_t)id * 3 * M + (size_t)code * M;
                for (uint32_t q = 0; q < M; ++q) {
                    const int c = reverse ? ord[q] : ord[M - 1 - q];
                    if (!isnan(dist[c])) {
                        if (sp >= ORC_STACK_CAP) {
                            dg |= ORC_DIAG_STACKCAP;
                            goto done;
                        }
                        stack[sp++] = b->base[id] + (uint32_t)c;
                        if (sp > 48) dg |= ORC_DIAG_STACK48;
                        dist[c] = NAN;
                    }
                }
            }
        }
    }
done:
    if (diag) *diag |= dg;
    return result;
}

/* FastNodeManager::intersectSaveDistance (fastNodeManager.h:1232-1326):
 * entries carry the child's t0; pops with t0 > tMax are skipped uncounted;
 * children are pushed whether or not the node test reported a hit. */
static int save_distance_impl(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index, int32_t* tri_index,
                              orc_counts* counts, uint32_t* diag, orc_seq* q) {
    uint32_t stack[ORC_STACK_CAP];
    float sdist[ORC_STACK_CAP];
    int sp = 1, result = 0;
    uint32_t dg = 0;
    float dist[16];
    float* ri = (float*)ray;
    const uint32_t M = b->M;
    stack[0] = 0;
    sdist[0] = 0.0f;
    for (uint32_t i = 0; i < M; ++i) dist[i] = NAN;
    const int code = max_abs_dim(ray->dx, ray->dy, ray->dz);
    const float dcode = code == 0 ? ray->dx : (code == 1 ? ray->dy : ray->dz);
    const int reverse = dcode <= 0.0f;
    while (sp != 0) {
        --sp;
        const uint32_t id = stack[sp];
        if (sdist[sp] > ray->tmax) continue;
        if (id >= b->node_count) {
            dg |= ORC_DIAG_BOGUS;
            continue;
        }
        if (is_leaf(b, id)) {
            const int k = orc_leaf_loop_count(b, b->base[id]);
            count_add(counts, q, id, 1, k);
            const int r = ORC_LEAF_TEST(b->tri, b->base[id], k, ri, (int)b->G);
            if (r != -1) {
                result = 1;
                *leaf_index = b->base[id];
                *tri_index = r;
            }
        } else {
            const int k = orc_node_loop_count(b, id);
            count_add(counts, q, id, 0, k);
            const float* bd = b->bounds + (size_t)id * 6 * M;
            (void)ORC_NODE_TEST(bd, k, dist, ri);
            const int8_t* ord = b->order + (size_t)id * 3 * M + (size_t)code * M;
            for (uint32_t q = 0; q < M; ++q) {
                const int c = reverse ? ord[q] : ord[M - 1 - q];
                if (!isnan(dist[c])) {
                    if (sp >= ORC_STACK_CAP) {
                        dg |= ORC_DIAG_STACKCAP;
                        goto done;
                    }
                    stack[sp] = b->base[id] + (uint32_t)c;
                    sdist[sp] = dist[c];
                    ++sp;
                    if (sp > 48) dg |= ORC_DIAG_STACK48;
                    dist[c] = NAN;
                }
            }
        }
    }
done:
    if (diag) *diag |= dg;
    return result;
}

/* FastNodeManager::intersectSecondary (fastNodeManager.h:1420-1493):
 * any-hit; children pushed in slot order 0..M-1; stops at the first leaf
 * whose triAnyHit is true. */
static int secondary_impl(const orc_bvh* b, orc_ray* ray, orc_counts* counts, uint32_t* diag, orc_seq* q) {
    uint32_t stack[ORC_STACK_CAP];
    int sp = 1;
    uint32_t dg = 0;
    float dist[16];
    const float* ri = (const float*)ray;
    const uint32_t M = b->M;
    stack[0] = 0;
    for (uint32_t i = 0; i < M; ++i) dist[i] = NAN;
    int result = 0;
    while (sp != 0) {
        const uint32_t id = stack[--sp];
        if (id >= b->node_count) {
            dg |= ORC_DIAG_BOGUS;
            continue;
        }
        if (is_leaf(b, id)) {
            const int k = orc_leaf_loop_count(b, b->base[id]);
            count_add(counts, q, id, 1, k);
            if (ORC_LEAF_ANY(b->tri, b->base[id], k, ri)) {
                result = 1;
                break;
            }
        } else {
            const int k = orc_node_loop_count(b, id);
            count_add(counts, q, id, 0, k);
            const float* bd = b->bounds + (size_t)id * 6 * M;
            if (ORC_NODE_TEST(bd, k, dist, ri)) {
                for (uint32_t i = 0; i < M; ++i) {
                    if (!isnan(dist[i])) {
                        if (sp >= ORC_STACK_CAP) {
                            dg |= ORC_DIAG_STACKCAP;
                            goto done;
                        }
                        stack[sp++] = b->base[id] + i;
                        if (sp > 48) dg |= ORC_DIAG_STACK48;
                        dist[i] = NAN;
                    }
                }
            }
        }
    }
done:
    if (diag) *diag |= dg;
    return result;
}

int orc_intersect(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index, int32_t* tri_index,
                  orc_counts* counts, uint32_t* diag) {
    return intersect_impl(b, ray, leaf_index, tri_index, counts, diag, NULL);
}

int orc_intersect_save_distance(const orc_bvh* b, orc_ray* ray, uint32_t* leaf_index,
                                int32_t* tri_index, orc_counts* counts, uint32_t* diag) {
    return save_distance_impl(b, ray, leaf_index, tri_index, counts, diag, NULL);
}

int orc_intersect_secondary(const orc_bvh* b, orc_ray* ray, orc_counts* counts, uint32_t* diag) {
    return secondary_impl(b, ray, counts, diag, NULL);
}

/* Visit sequence of each ray under the per-ray traversal of `mode`: the
 * order in which a lockstep wide traversal (compactNode.cpp:358-601, one
 * visit per ray per step) takes the ray's visits.  ids: n x cap; lens[i] =
 * visits of ray i (ids keeps the first cap). */
void orc_visit_sequences(const orc_bvh* b, const orc_ray* rays, uint32_t n, int mode, uint32_t* ids,
                         uint32_t cap, uint32_t* lens) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; ++i) {
        orc_ray r = rays[i];
        orc_seq q = {ids + (size_t)i * cap, cap, 0};
        uint32_t leaf = 0;
        int32_t slot = -1;
        if (mode == ORC_ANY) (void)secondary_impl(b, &r, NULL, NULL, &q);
        else if (mode == ORC_SAVE_DIST) (void)save_distance_impl(b, &r, &leaf, &slot, NULL, NULL, &q);
        else (void)intersect_impl(b, &r, &leaf, &slot, NULL, NULL, &q);
        lens[i] = q.len;
    }
}

/* Stream traversal state shared by the two wide forms. */
typedef struct wide_state {
    uint32_t count;
    int32_t* stack;    /* count * ORC_STACK_CAP */
    int* sp;
    uint16_t* work[2];
    float* dist;       /* count * 16 (per ray) or 16 (shared) */
} wide_state;

static int wide_alloc(wide_state* s, uint32_t count) {
    s->count = count;
    s->stack = (int32_t*)malloc(sizeof(int32_t) * ORC_STACK_CAP * count);
    s->sp = (int*)malloc(sizeof(int) * count);
    s->work[0] = (uint16_t*)malloc(sizeof(uint16_t) * count);
    s->work[1] = (uint16_t*)malloc(sizeof(uint16_t) * count);
    s->dist = (float*)malloc(sizeof(float) * 16 * count);
    if (!s->stack || !s->sp || !s->work[0] || !s->work[1] || !s->dist) return 0;
    for (uint32_t i = 0; i < 16 * count; ++i) s->dist[i] = NAN;
    return 1;
}

static void wide_free(wide_state* s) {
    free(s->stack);
    free(s->sp);
    free(s->work[0]);
    free(s->work[1]);
    free(s->dist);
}

/* node entry of the stream stack: child type from childType (inner) or, with
 * per-ray semantics, from the node itself so that any pushed id resolves
 * like the per-ray isnan(bounds[0]) test. */
static inline int32_t wide_entry(const orc_bvh* b, uint32_t id, int child_is_inner, int per_ray) {
    int inner = per_ray ? (id < b->node_count ? !is_leaf(b, id) : 1) : child_is_inner;
    return inner ? (int32_t)id : -(int32_t)id;
}

/* FastNodeManager::intersectWideAlternative (fastNodeManager.h:835-1049).
 * Without flags every ray follows exactly its per-ray DFS (intersect). */
void orc_intersect_wide(const orc_bvh* b, orc_ray* rays, uint32_t count, uint32_t flags,
                        uint32_t* leaf_index, int32_t* tri_index, orc_counts* counts, uint32_t* diag) {
    wide_state s;
    const uint32_t M = b->M;
    const int per_ray = !(flags & (ORC_WIDE_AXIS_BUG | ORC_WIDE_SHARED_DIST));
    if (!wide_alloc(&s, count)) {
        wide_free(&s);
        return;
    }
    uint8_t* major = (uint8_t*)malloc(count);
    for (uint32_t i = 0; i < count; ++i) {
        tri_index[i] = -1;
        s.sp[i] = 1;
        const int root_inner = (flags & ORC_WIDE_ROOT_AS_NODE) ? 1 : !is_leaf(b, 0);
        s.stack[(size_t)i * ORC_STACK_CAP] = root_inner ? 0 : -0;
        s.work[0][i] = (uint16_t)i;
        const int code = max_abs_dim(rays[i].dx, rays[i].dy, rays[i].dz);
        const float dc = code == 0 ? rays[i].dx : (code == 1 ? rays[i].dy : rays[i].dz);
        major[i] = (uint8_t)((code << 1) + (dc <= 0.0f ? 1 : 0));
    }
    /* a root leaf is entry -0 == 0: route it through the leaf list instead */
    int rootLeaf = !(flags & ORC_WIDE_ROOT_AS_NODE) && is_leaf(b, 0);
    uint16_t* cur = s.work[0];
    uint16_t* nxt = s.work[1];
    uint32_t nodeRays = rootLeaf ? 0 : count, leafRays = 0;
    if (rootLeaf) {
        for (uint32_t i = 0; i < count; ++i) cur[count - 1 - i] = (uint16_t)i;
        leafRays = count;
    }
    while (nodeRays != 0 || leafRays != 0) {
        uint32_t nodeNext = 0, leafNext = 0;
        for (uint32_t i = 0; i < nodeRays; ++i) {
            const uint16_t r = cur[i];
            float* ri = (float*)&rays[r];
            int32_t* st = s.stack + (size_t)r * ORC_STACK_CAP;
            const uint32_t id = (uint32_t)st[--s.sp[r]];
            float* dist = (flags & ORC_WIDE_SHARED_DIST) ? s.dist : s.dist + (size_t)r * 16;
            if (id >= b->node_count) {
                if (diag) diag[r] |= ORC_DIAG_BOGUS;
            } else {
                const int k = orc_node_loop_count(b, id);
                if (counts) count_add(&counts[r], NULL, 0, 0, k);
                if (ORC_NODE_TEST(b->bounds + (size_t)id * 6 * M, k, dist, ri)) {
                    const uint32_t axis = (flags & ORC_WIDE_AXIS_BUG) ? (major[i] >> 1) : (major[r] >> 1);
                    const int8_t* ord = b->order + (size_t)id * 3 * M + (size_t)axis * M;
                    const int fwd = major[r] & 1;
                    for (uint32_t q = 0; q < M; ++q) {
                        const int c = fwd ? ord[q] : ord[M - 1 - q];
                        if (!isnan(dist[c])) {
                            if (s.sp[r] >= ORC_STACK_CAP) {
                                if (diag) diag[r] |= ORC_DIAG_STACKCAP;
                                s.sp[r] = 0;
                                break;
                            }
                            st[s.sp[r]++] = wide_entry(b, b->base[id] + (uint32_t)c,
                                                       (b->child_type[id] >> c) & 1, per_ray);
                            if (s.sp[r] > 48 && diag) diag[r] |= ORC_DIAG_STACK48;
                            dist[c] = NAN;
                        }
                    }
                }
            }
            if (s.sp[r] != 0) {
                if (st[s.sp[r] - 1] >= 0) nxt[nodeNext++] = r;
                else nxt[count - 1 - (leafNext++)] = r;
            }
        }
        for (uint32_t i = 0; i < leafRays; ++i) {
            const uint16_t r = cur[count - 1 - i];
            float* ri = (float*)&rays[r];
            int32_t* st = s.stack + (size_t)r * ORC_STACK_CAP;
            const uint32_t id = (uint32_t)(-st[--s.sp[r]]);
            if (id >= b->node_count) {
                if (diag) diag[r] |= ORC_DIAG_BOGUS;
            } else {
                const int k = orc_leaf_loop_count(b, b->base[id]);
                if (counts) count_add(&counts[r], NULL, 0, 1, k);
                const int res = ORC_LEAF_TEST(b->tri, b->base[id], k, ri, (int)b->G);
                if (res != -1) {
                    leaf_index[r] = b->base[id];
                    tri_index[r] = res;
                }
            }
            if (s.sp[r] != 0) {
                if (st[s.sp[r] - 1] >= 0) nxt[nodeNext++] = r;
                else nxt[count - 1 - (leafNext++)] = r;
            }
        }
        leafRays = leafNext;
        nodeRays = nodeNext;
        uint16_t* t = cur;
        cur = nxt;
        nxt = t;
    }
    free(major);
    wide_free(&s);
}

/* FastNodeManager::intersectSecondaryWideAlternative (fastNodeManager.h:1051-1230):
 * rays whose origin x is NaN (primary miss) are skipped; result[r]++ on hit. */
void orc_intersect_secondary_wide(const orc_bvh* b, orc_ray* rays, uint32_t count, uint32_t flags,
                                  uint16_t* result, orc_counts* counts, uint32_t* diag) {
    wide_state s;
    const uint32_t M = b->M;
    const int per_ray = !(flags & (ORC_WIDE_AXIS_BUG | ORC_WIDE_SHARED_DIST));
    if (!wide_alloc(&s, count)) {
        wide_free(&s);
        return;
    }
    uint16_t* cur = s.work[0];
    uint16_t* nxt = s.work[1];
    uint32_t nodeRays = 0, leafRays = 0;
    const int rootLeaf = !(flags & ORC_WIDE_ROOT_AS_NODE) && is_leaf(b, 0);
    for (uint32_t i = 0; i < count; ++i) {
        s.sp[i] = 1;
        s.stack[(size_t)i * ORC_STACK_CAP] = 0;
        if (!isnan(rays[i].ox)) {
            if (rootLeaf) cur[count - 1 - (leafRays++)] = (uint16_t)i;
            else cur[nodeRays++] = (uint16_t)i;
        }
    }
    while (nodeRays != 0 || leafRays != 0) {
        uint32_t nodeNext = 0, leafNext = 0;
        for (uint32_t i = 0; i < nodeRays; ++i) {
            const uint16_t r = cur[i];
            float* ri = (float*)&rays[r];
            int32_t* st = s.stack + (size_t)r * ORC_STACK_CAP;
            const uint32_t id = (uint32_t)st[--s.sp[r]];
            float* dist = (flags & ORC_WIDE_SHARED_DIST) ? s.dist : s.dist + (size_t)r * 16;
            if (id >= b->node_count) {
                if (diag) diag[r] |= ORC_DIAG_BOGUS;
            } else {
                const int k = orc_node_loop_count(b, id);
                if (counts) count_add(&counts[r], NULL, 0, 0, k);
                if (ORC_NODE_TEST(b->bounds + (size_t)id * 6 * M, k, dist, ri)) {
                    for (uint32_t c = 0; c < M; ++c) {
                        if (!isnan(dist[c])) {
                            if (s.sp[r] >= ORC_STACK_CAP) {
                                if (diag) diag[r] |= ORC_DIAG_STACKCAP;
                                s.sp[r] = 0;
                                break;
                            }
                            st[s.sp[r]++] = wide_entry(b, b->base[id] + c, (b->child_type[id] >> c) & 1, per_ray);
                            if (s.sp[r] > 48 && diag) diag[r] |= ORC_DIAG_STACK48;
                            dist[c] = NAN;
                        }
                    }
                }
            }
            if (s.sp[r] != 0) {
                if (st[s.sp[r] - 1] >= 0) nxt[nodeNext++] = r;
                else nxt[count - 1 - (leafNext++)] = r;
            }
        }
        for (uint32_t i = 0; i < leafRays; ++i) {
            const uint16_t r = cur[count - 1 - i];
            float* ri = (float*)&rays[r];
            int32_t* st = s.stack + (size_t)r * ORC_STACK_CAP;
            const uint32_t id = (uint32_t)(-st[--s.sp[r]]);
            int hit = 0;
            if (id >= b->node_count) {
                if (diag) diag[r] |= ORC_DIAG_BOGUS;
            } else {
                const int k = orc_leaf_loop_count(b, b->base[id]);
                if (counts) count_add(&counts[r], NULL, 0, 1, k);
                hit = ORC_LEAF_ANY(b->tri, b->base[id], k, ri);
            }
            if (hit) {
                result[r]++;
                continue;
            }
            if (s.sp[r] != 0) {
                if (st[s.sp[r] - 1] >= 0) nxt[nodeNext++] = r;
                else nxt[count - 1 - (leafNext++)] = r;
            }
        }
        leafRays = leafNext;
        nodeRays = nodeNext;
        uint16_t* t = cur;
        cur = nxt;
        nxt = t;
    }
    wide_free(&s);
}

static void fill_hit(const orc_bvh* b, const orc_ray* orig, float tfinal, int found,
                     uint32_t leaf, int32_t slot, orc_hit* h) {
    if (!found || slot < 0) {
        h->slot = -1;
        h->leaf = 0;
        h->t = tfinal;
        h->b1 = h->b2 = 0.0f;
        return;
    }
    const int k = orc_leaf_loop_count(b, leaf);
    float p0[3], p1[3], p2[3], t, b1, b2;
    load_tri(b->tri, leaf, k, slot, p0, p1, p2);
    orc_ray r = *orig;
    r.tmax = tfinal;
    (void)mt32(p0, p1, p2, (const float*)&r, &t, &b1, &b2);
    h->slot = slot;
    h->leaf = leaf;
    h->t = tfinal;
    h->b1 = b1;
    h->b2 = b2;
}

int orc_trace(const orc_bvh* b, const orc_ray* rays, uint32_t n, int mode, int stream,
              uint32_t tile, int threads, orc_hit* hits, uint8_t* occluded, orc_counts* counts,
              uint32_t* diag) {
    if (tile == 0) tile = 256;
    if (tile > 4096) tile = 4096;
    const long tiles = ((long)n + tile - 1) / tile;
    int used = 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
        orc_ray* local = (orc_ray*)malloc(sizeof(orc_ray) * tile);
        uint32_t* li = (uint32_t*)malloc(sizeof(uint32_t) * tile);
        int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * tile);
        uint16_t* res = (uint16_t*)malloc(sizeof(uint16_t) * tile);
#pragma omp for schedule(dynamic, 4)
        for (long t = 0; t < tiles; ++t) {
            const uint32_t first = (uint32_t)(t * tile);
            const uint32_t cnt = (n - first < tile) ? (n - first) : tile;
            memcpy(local, rays + first, sizeof(orc_ray) * cnt);
            if (counts) memset(counts + first, 0, sizeof(orc_counts) * cnt);
            if (diag) memset(diag + first, 0, sizeof(uint32_t) * cnt);
            if (stream && mode == ORC_CLOSEST) {
                orc_intersect_wide(b, local, cnt, 0, li, ti, counts ? counts + first : NULL,
                                   diag ? diag + first : NULL);
                for (uint32_t j = 0; j < cnt; ++j)
                    fill_hit(b, &rays[first + j], local[j].tmax, ti[j] >= 0, li[j], ti[j], &hits[first + j]);
            } else if (stream && mode == ORC_ANY) {
                memset(res, 0, sizeof(uint16_t) * cnt);
                orc_intersect_secondary_wide(b, local, cnt, 0, res, counts ? counts + first : NULL,
                                             diag ? diag + first : NULL);
                for (uint32_t j = 0; j < cnt; ++j) occluded[first + j] = res[j] ? 1 : 0;
            } else {
                for (uint32_t j = 0; j < cnt; ++j) {
                    orc_ray r = local[j];
                    orc_counts* cc = counts ? &counts[first + j] : NULL;
                    uint32_t* dd = diag ? &diag[first + j] : NULL;
                    if (mode == ORC_ANY) {
                        occluded[first + j] = (uint8_t)orc_intersect_secondary(b, &r, cc, dd);
                    } else {
                        uint32_t leaf = 0;
                        int32_t slot = -1;
                        int hit = (mode == ORC_SAVE_DIST)
                                      ? orc_intersect_save_distance(b, &r, &leaf, &slot, cc, dd)
                                      : orc_intersect(b, &r, &leaf, &slot, cc, dd);
                        fill_hit(b, &rays[first + j], r.tmax, hit, leaf, slot, &hits[first + j]);
                    }
                }
            }
        }
        free(local);
        free(li);
        free(ti);
        free(res);
    }
    return used;
}

int orc_mt_f64(const orc_ray* ray, const float* p, double* tout, double* b1out, double* b2out) {
    const double ox = ray->ox, oy = ray->oy, oz = ray->oz, dx = ray->dx, dy = ray->dy, dz = ray->dz;
    const double e1x = (double)p[3] - p[0], e1y = (double)p[4] - p[1], e1z = (double)p[5] - p[2];
    const double e2x = (double)p[6] - p[0], e2y = (double)p[7] - p[1], e2z = (double)p[8] - p[2];
    const double s1x = dy * e2z - dz * e2y, s1y = dz * e2x - dx * e2z, s1z = dx * e2y - dy * e2x;
    const double div = s1x * e1x + s1y * e1y + s1z * e1z;
    if (div == 0.0) return 0;
    const double ddx = ox - p[0], ddy = oy - p[1], ddz = oz - p[2];
    const double b1 = (ddx * s1x + ddy * s1y + ddz * s1z) / div;
    const double s2x = ddy * e1z - ddz * e1y, s2y = ddz * e1x - ddx * e1z, s2z = ddx * e1y - ddy * e1x;
    const double b2 = (dx * s2x + dy * s2y + dz * s2z) / div;
    const double t = (e2x * s2x + e2y * s2y + e2z * s2z) / div;
    *tout = t;
    *b1out = b1;
    *b2out = b2;
    return (b1 >= 0.0 && b2 >= 0.0 && b1 + b2 <= 1.0 && t >= 0.0);
}

int32_t orc_bruteforce_closest(const float* tris, uint32_t n, const orc_ray* ray, float* tout) {
    float ri[10];
    memcpy(ri, ray, sizeof(ri));
    int32_t best = -1;
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = tris + (size_t)i * 9;
        float t, b1, b2;
        /* ri[0] holds the best t so far: equal t passes the test, the first
         * triangle reaching the minimum is kept */
        if (mt32(p, p + 3, p + 6, ri, &t, &b1, &b2) && t == t && (best < 0 || t < ri[0])) {
            ri[0] = t;
            best = (int32_t)i;
        }
    }
    *tout = ri[0];
    return best;
}

int orc_bruteforce_any(const float* tris, uint32_t n, const orc_ray* ray) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = tris + (size_t)i * 9;
        float t, b1, b2;
        if (mt32(p, p + 3, p + 6, (const float*)ray, &t, &b1, &b2)) return 1;
    }
    return 0;
}
