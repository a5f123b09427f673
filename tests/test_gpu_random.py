"""Randomised parity on triangle soups the scene generators never produce:
degenerate (repeated vertex, collinear) and duplicated triangles, triangles
in axis-aligned planes, coordinates snapped to a grid (many equal box
planes and sort keys), scales from 1e-3 to 1e20 (products that overflow to
infinity), built at random node/leaf size, gang size and builder options,
and traced with rays aimed at vertices, edge midpoints and duplicated
triangles, axis-parallel rays, rays with zero direction components, origins
on grid planes and short tMax.  Every mode and schedule the boundary
offers (closest hit per-ray / ray-stream / cooperative, save distance,
any hit per-ray / ray-stream / cooperative) must equal the oracle bit for
bit: hit identity, t, b1, b2, occlusion and the node/leaf counts.  Any-hit
entries that point past the node array (the reference's undefined
behaviour) are skipped by both sides; the GPU counts those rays as guarded
errors and the count must equal the oracle's flagged rays.

A build without sortEachSplit leaves every order row zero, as the reference
does (aabb.cpp:191-195 computes no order; the FastNode constructor pads the
empty rows with 0, fastNodeManager.h:65-79).  The closest-hit traversals,
which walk the order row, then only ever push child 0, while the any-hit
traversal pushes in slot order; those cases find few closest hits by
design."""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu

SEEDS = list(range(16))


def _soup(rng, seed):
    n = int(rng.integers(40, 2500))
    scale = 1e20 if seed == 7 else float(10.0 ** rng.uniform(-3, 4))
    centres = rng.uniform(-1.0, 1.0, size=(n, 1, 3)) * scale
    size = scale * float(10.0 ** rng.uniform(-2.5, -0.7))
    v = centres + rng.normal(size=(n, 3, 3)) * size
    u = rng.random(n)
    rep = np.nonzero(u < 0.08)[0]
    v[rep, 1] = v[rep, 0]                                   # repeated vertex
    col = np.nonzero((u >= 0.08) & (u < 0.13))[0]
    v[col, 2] = v[col, 0] + 2.0 * (v[col, 1] - v[col, 0])   # collinear
    flat = np.nonzero((u >= 0.13) & (u < 0.25))[0]
    ax = rng.integers(0, 3, flat.size)
    v[flat, :, ax] = v[flat, 0, ax][:, None]                # axis-aligned plane
    snap = np.nonzero((u >= 0.25) & (u < 0.45))[0]
    step = size * 0.5
    v[snap] = np.round(v[snap] / step) * step               # grid-snapped
    tris = v.reshape(n, 9).astype(np.float32)
    dup = rng.choice(n, max(1, n // 20), replace=False)
    return np.concatenate([tris, tris[dup]]), step


def _rays(rt, rng, tris, step, m=6000):
    v = tris.reshape(-1, 3, 3).astype(np.float64)
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    span = np.maximum(hi - lo, 1e-30)
    o = lo - 0.2 * span + rng.random((m, 3)) * 1.4 * span
    t = lo + rng.random((m, 3)) * span
    k = rng.integers(0, 6, m)
    pick = rng.integers(0, len(v), m)
    t[k == 1] = v[pick[k == 1], 0]                           # at a vertex
    t[k == 2] = 0.5 * (v[pick[k == 2], 0] + v[pick[k == 2], 1])  # at an edge midpoint
    t[k == 3] = v[pick[k == 3]].mean(1)                      # at a centroid (duplicates tie)
    par = np.nonzero(k == 4)[0]                              # axis-parallel
    axes = rng.integers(0, 3, par.size)
    t[par] = o[par]
    t[par, axes] += np.where(rng.random(par.size) < 0.5, -1.0, 1.0) * span[axes]
    zero = np.nonzero(k == 5)[0]                             # one zero component, origin on a grid plane
    t[zero, 1] = o[zero, 1]
    o[zero, 0] = np.round(o[zero, 0] / step) * step
    rays = rt.make_rays(o.astype(np.float32), t.astype(np.float32))
    short = rng.random(m) < 0.2
    rays["tmax"][short] = rng.uniform(0.0, 1.5, short.sum()).astype(np.float32)
    return rays


def _config(rng):
    N = int(rng.choice([4, 8, 12, 16]))
    L = int(rng.choice([4, 8, 12, 16]))
    G = int(rng.choice([4, 8, 16]))
    return dict(branching=N, leaf_size=L, gang=G, sort_each_split=bool(rng.random() < 0.7),
                leaf_split_option=int(rng.integers(0, 2)), sah_factor=float(rng.uniform(0.8, 1.2)))


def _errors(scene):
    return scene.stats()["error_rays"]


@pytest.mark.parametrize("seed", SEEDS)
def test_random_soup(rt, seed):
    rng = np.random.default_rng(7000 + seed)
    tris, step = _soup(rng, seed)
    cfg = _config(rng)
    print(f"random soup seed {seed}: {len(tris)} tris, {cfg}", flush=True)
    bvh = rt.Bvh(tris, cfg["branching"], cfg["leaf_size"], gang=cfg["gang"],
                 sort_each_split=cfg["sort_each_split"], leaf_split_option=cfg["leaf_split_option"],
                 sah_factor=cfg["sah_factor"])
    rays = _rays(rt, rng, tris, step)
    scene = rt.Scene(bvh)
    o = oracle_for(bvh)
    try:
        ref = o.trace(rays)
        ref_sd = o.trace(rays, mode=oracle.ORC_SAVE_DIST)
        ref_any = o.trace(rays, mode=oracle.ORC_ANY)
        bogus = int(((ref_any["diag"] & oracle.DIAG_BOGUS) != 0).sum())
        for kw in ({}, {"stream": True}, {"coop": True}):
            print(f"  closest {kw}", flush=True)
            try:
                h, c = scene.intersect(rays, counts=True, **kw)
            except rt.RtError as e:
                # the cooperative kernel keeps a ray's whole stack in LDS
                assert kw == {"coop": True} and "too deep" in str(e), e
                continue
            assert_closest_equal(h, ref["hits"], bvh, c, ref["counts"], f"seed {seed} closest {kw}")
        h, c = scene.intersect(rays, mode="save_distance", counts=True)
        assert_closest_equal(h, ref_sd["hits"], bvh, c, ref_sd["counts"], f"seed {seed} save distance")
        for kw in ({}, {"stream": True}, {"coop": True}):
            print(f"  any {kw}", flush=True)
            e0 = _errors(scene)
            try:
                occ, c = scene.intersect(rays, mode="any", counts=True, **kw)
            except rt.RtError as e:
                assert kw == {"coop": True} and "too deep" in str(e), e
                continue
            assert np.array_equal(occ, ref_any["occluded"]), f"seed {seed} any {kw}"
            for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
                assert np.array_equal(c[f], ref_any["counts"][f]), f"seed {seed} any {kw} {f}"
            assert _errors(scene) - e0 == bogus, f"seed {seed} any {kw}: guarded rays"
        # closest-hit traversals never push a padding slot
        assert ((ref["diag"] & oracle.DIAG_BOGUS) == 0).all()
    finally:
        scene.close()
