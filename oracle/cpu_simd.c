/*
 * ORACLE — test infrastructure and CPU baseline, not product code.
 *
 * The oracle's traversals (cpu_ref.c, included below unchanged) with the node
 * and leaf tests in the shape the reference runs them: 4-wide SSE code, one
 * child slot or triangle per lane, k slots four at a time, as the ISPC
 * kernels compile for the sse4-i32x4 target (compileISPC.cmd:2;
 * rayTracerTemplate.ispc:25-234: foreach over the k slots, maxps/minps for
 * max()/min(), divps for the reciprocal, reduce_min/reduce_max over the
 * program instances).  The arithmetic is the same FP32 sequence as the
 * scalar tests, so every result equals cpu_ref.c's bit for bit.  Leaf tests
 * of gangs other than 4 run the scalar test.
 * Build: gcc -O3 -msse4.1 -ffp-contract=off -fno-fast-math -fopenmp (Makefile).
 */
#include <immintrin.h>
#include <stdint.h>

#include "cpu_ref.h"

int orc_aabb_intersect(const float* bounds, int k, float* dist, const float* ri);
int orc_tri_intersect(const float* tri, uint32_t tri_id, int k, float* ri, int G);
int orc_tri_anyhit(const float* tri, uint32_t tri_id, int k, const float* ri);

/* aabbIntersect{k} (rayTracerTemplate.ispc:170-234), four slots per step. */
static int simd_aabb_intersect(const float* bounds, int k, float* dist, const float* ri) {
    const __m128 zero = _mm_setzero_ps(), tmax = _mm_set1_ps(ri[0]);
    const __m128 ox = _mm_set1_ps(ri[1]), oy = _mm_set1_ps(ri[2]), oz = _mm_set1_ps(ri[3]);
    const __m128 ix = _mm_set1_ps(ri[4]), iy = _mm_set1_ps(ri[5]), iz = _mm_set1_ps(ri[6]);
    int any = 0;
    for (int j = 0; j < k; j += 4) {
        const __m128 tnx = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + j), ox), ix);
        const __m128 tny = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + k + j), oy), iy);
        const __m128 tnz = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + 2 * k + j), oz), iz);
        const __m128 tfx = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + 3 * k + j), ox), ix);
        const __m128 tfy = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + 4 * k + j), oy), iy);
        const __m128 tfz = _mm_mul_ps(_mm_sub_ps(_mm_loadu_ps(bounds + 5 * k + j), oz), iz);
        /* per axis: swap where near > far, then t0 = max(near, t0),
         * t1 = min(far, t1); _mm_max_ps(a, b) = (a > b) ? a : b */
        __m128 m = _mm_cmpgt_ps(tnx, tfx);
        __m128 t0 = _mm_max_ps(_mm_blendv_ps(tnx, tfx, m), zero);
        __m128 t1 = _mm_min_ps(_mm_blendv_ps(tfx, tnx, m), tmax);
        m = _mm_cmpgt_ps(tny, tfy);
        t0 = _mm_max_ps(_mm_blendv_ps(tny, tfy, m), t0);
        t1 = _mm_min_ps(_mm_blendv_ps(tfy, tny, m), t1);
        m = _mm_cmpgt_ps(tnz, tfz);
        t0 = _mm_max_ps(_mm_blendv_ps(tnz, tfz, m), t0);
        t1 = _mm_min_ps(_mm_blendv_ps(tfz, tnz, m), t1);
        const __m128 hit = _mm_cmple_ps(t0, t1);
        if (_mm_movemask_ps(hit)) {
            any = 1;
            _mm_storeu_ps(dist + j, _mm_blendv_ps(_mm_loadu_ps(dist + j), t0, hit));
        }
    }
    return any;
}

/* Moeller-Trumbore on four triangles (rayTracerTemplate.ispc:49-75): the hit
 * mask, with t in *tout. */
static inline __m128 simd_mt(const float* b, int k, int j, const float* ri, __m128* tout) {
    const __m128 ox = _mm_set1_ps(ri[1]), oy = _mm_set1_ps(ri[2]), oz = _mm_set1_ps(ri[3]);
    const __m128 dx = _mm_set1_ps(ri[7]), dy = _mm_set1_ps(ri[8]), dz = _mm_set1_ps(ri[9]);
    const __m128 p0x = _mm_loadu_ps(b + j), p0y = _mm_loadu_ps(b + k + j), p0z = _mm_loadu_ps(b + 2 * k + j);
    const __m128 e1x = _mm_sub_ps(_mm_loadu_ps(b + 3 * k + j), p0x);
    const __m128 e1y = _mm_sub_ps(_mm_loadu_ps(b + 4 * k + j), p0y);
    const __m128 e1z = _mm_sub_ps(_mm_loadu_ps(b + 5 * k + j), p0z);
    const __m128 e2x = _mm_sub_ps(_mm_loadu_ps(b + 6 * k + j), p0x);
    const __m128 e2y = _mm_sub_ps(_mm_loadu_ps(b + 7 * k + j), p0y);
    const __m128 e2z = _mm_sub_ps(_mm_loadu_ps(b + 8 * k + j), p0z);
    /* s1 = cross(dir, e2) */
    const __m128 s1x = _mm_sub_ps(_mm_mul_ps(dy, e2z), _mm_mul_ps(dz, e2y));
    const __m128 s1y = _mm_sub_ps(_mm_mul_ps(dz, e2x), _mm_mul_ps(dx, e2z));
    const __m128 s1z = _mm_sub_ps(_mm_mul_ps(dx, e2y), _mm_mul_ps(dy, e2x));
    const __m128 divisor =
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(s1x, e1x), _mm_mul_ps(s1y, e1y)), _mm_mul_ps(s1z, e1z));
    const __m128 zero = _mm_setzero_ps(), one = _mm_set1_ps(1.0f);
    __m128 hit = _mm_cmpneq_ps(divisor, zero); /* a NaN divisor keeps the hit */
    const __m128 inv = _mm_div_ps(one, divisor);
    const __m128 ddx = _mm_sub_ps(ox, p0x), ddy = _mm_sub_ps(oy, p0y), ddz = _mm_sub_ps(oz, p0z);
    const __m128 b1 = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(ddx, s1x), _mm_mul_ps(ddy, s1y)), _mm_mul_ps(ddz, s1z)), inv);
    hit = _mm_andnot_ps(_mm_or_ps(_mm_cmplt_ps(b1, zero), _mm_cmpgt_ps(b1, one)), hit);
    /* s2 = cross(d, e1) */
    const __m128 s2x = _mm_sub_ps(_mm_mul_ps(ddy, e1z), _mm_mul_ps(ddz, e1y));
    const __m128 s2y = _mm_sub_ps(_mm_mul_ps(ddz, e1x), _mm_mul_ps(ddx, e1z));
    const __m128 s2z = _mm_sub_ps(_mm_mul_ps(ddx, e1y), _mm_mul_ps(ddy, e1x));
    const __m128 b2 = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(dx, s2x), _mm_mul_ps(dy, s2y)), _mm_mul_ps(dz, s2z)), inv);
    hit = _mm_andnot_ps(_mm_or_ps(_mm_cmplt_ps(b2, zero), _mm_cmpgt_ps(_mm_add_ps(b1, b2), one)), hit);
    const __m128 t = _mm_mul_ps(
        _mm_add_ps(_mm_add_ps(_mm_mul_ps(e2x, s2x), _mm_mul_ps(e2y, s2y)), _mm_mul_ps(e2z, s2z)), inv);
    hit = _mm_andnot_ps(_mm_or_ps(_mm_cmplt_ps(t, zero), _mm_cmpgt_ps(t, _mm_set1_ps(ri[0]))), hit);
    *tout = t;
    return hit;
}

/* triIntersect{k} (rayTracerTemplate.ispc:25-108), gang 4: each lane keeps
 * its first strict minimum (tStore from 1e6), bestT = reduce_min, the result
 * is reduce_max of the slot index over lanes whose tStore == bestT. */
static int simd_tri_intersect(const float* tri, uint32_t tri_id, int k, float* ri, int G) {
    if (G != 4) return orc_tri_intersect(tri, tri_id, k, ri, G);
    const float* b = tri + tri_id;
    __m128 tStore = _mm_set1_ps(1000000.0f);
    __m128i orig = _mm_set1_epi32(-1);
    const __m128i lane = _mm_setr_epi32(0, 1, 2, 3);
    int any = 0;
    for (int j = 0; j < k; j += 4) {
        __m128 t;
        const __m128 hit = simd_mt(b, k, j, ri, &t);
        const __m128 upd = _mm_and_ps(hit, _mm_cmpgt_ps(tStore, t));
        if (_mm_movemask_ps(upd)) {
            any = 1;
            tStore = _mm_blendv_ps(tStore, t, upd);
            orig = _mm_castps_si128(_mm_blendv_ps(_mm_castsi128_ps(orig),
                                                  _mm_castsi128_ps(_mm_add_epi32(lane, _mm_set1_epi32(j))), upd));
        }
    }
    if (!any) return -1;
    float ts[4];
    int32_t os[4];
    _mm_storeu_ps(ts, tStore);
    _mm_storeu_si128((__m128i*)os, orig);
    float bestT = ts[0];
    for (int g = 1; g < 4; ++g) bestT = (ts[g] < bestT) ? ts[g] : bestT;
    ri[0] = bestT;
    int idx = -1;
    for (int g = 0; g < 4; ++g)
        if (ts[g] == bestT && os[g] > idx) idx = os[g];
    return idx;
}

/* triAnyHit{k} (rayTracerTemplate.ispc:111-168) */
static int simd_tri_anyhit(const float* tri, uint32_t tri_id, int k, const float* ri) {
    const float* b = tri + tri_id;
    for (int j = 0; j < k; j += 4) {
        __m128 t;
        if (_mm_movemask_ps(simd_mt(b, k, j, ri, &t))) return 1;
    }
    return 0;
}

#define ORC_NODE_TEST simd_aabb_intersect
#define ORC_LEAF_TEST simd_tri_intersect
#define ORC_LEAF_ANY simd_tri_anyhit
#include "cpu_ref.c"
