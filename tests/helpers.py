"""Shared helpers for the parity tests (the oracle is the checker)."""
import numpy as np

import oracle


def oracle_for(bvh):
    return oracle.Oracle.from_bvh(bvh)


def assert_closest_equal(gpu_hits, orc_hits, bvh, gpu_counts=None, orc_counts=None, label=""):
    """Bit-exact hit identity, t, barycentrics and counts."""
    g, o = gpu_hits, orc_hits
    bad = np.nonzero(g["slot"] != o["slot"])[0]
    assert bad.size == 0, f"{label}: slot differs for {bad.size} rays, first {bad[:5]}"
    hit = o["slot"] >= 0
    assert np.array_equal(g["leaf"][hit], o["leaf"][hit]), f"{label}: leaf index differs"
    tri_expect = np.where(hit, bvh.tri_ids[(o["leaf"] // (9 * bvh.leaf_memory)) * bvh.leaf_memory
                                           + np.maximum(o["slot"], 0)], -1)
    assert np.array_equal(g["tri"], tri_expect), f"{label}: triangle id differs"
    for f in ("t", "b1", "b2"):
        a, b = g[f][hit], o[f][hit]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
            f"{label}: {f} not bit-exact (max abs diff {np.max(np.abs(a - b)) if a.size else 0})"
    assert np.array_equal(g["t"][~hit].view(np.uint32), o["t"][~hit].view(np.uint32)), f"{label}: miss t"
    if gpu_counts is not None:
        for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
            d = np.nonzero(gpu_counts[f] != orc_counts[f])[0]
            assert d.size == 0, f"{label}: counts.{f} differ for {d.size} rays, first {d[:5]}"
