"""The config-1 regression fixture (tests/golden/config1_n4l4.npz, written by
tests/golden/make_config1.py): the builder's topology, the camera rays and
the oracle's results of the 256x256 frame are unchanged; on the GPU the
traversal reproduces the fixture's sampled rays bit for bit."""
import os

import numpy as np
import pytest

from tests.golden.make_config1 import digests, frame

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_n4l4.npz")


def load():
    z = np.load(FIXTURE, allow_pickle=False)
    return {k: z[k] for k in z.files}


def test_config1_fixture_cpu():
    g = load()
    tris, bvh, rays, closest = frame()
    assert bvh.node_count == int(g["node_count"]) and bvh.max_depth == int(g["max_depth"])
    d = dict(x.split("=", 1) for x in g["digests"])
    assert digests(tris, bvh, rays, closest) == d
    sel = g["ray_index"]
    assert np.array_equal(rays[sel].view(np.uint32), g["rays"].view(np.uint32))
    assert np.array_equal(closest["hits"][sel], g["hits"]) and np.array_equal(closest["counts"][sel], g["counts"])


@pytest.mark.gpu
def test_config1_fixture_gpu(rt):
    g = load()
    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 4, 4)
    rays = g["rays"].view(rt.RAY_DTYPE)
    print(f"golden config1: {len(rays)} rays", flush=True)
    hits, cnt = rt.Scene(bvh).intersect(rays, counts=True)
    ref = g["hits"]
    assert np.array_equal(hits["slot"], ref["slot"]) and np.array_equal(hits["leaf"], ref["leaf"])
    for f in ("t", "b1", "b2"):
        assert np.array_equal(hits[f].view(np.uint32), ref[f].view(np.uint32)), f
    for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
        assert np.array_equal(cnt[f], g["counts"][f]), f
