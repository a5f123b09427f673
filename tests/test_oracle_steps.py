"""Visit sequences and per-step tile records of the oracle (the counting of
CompactNodeManager::intersectWideAlternative, compactNode.cpp:358-601).

Hand-built trees pin the visit order (worked out by hand below from the
order rows and the reference's push rules); on generated scenes the
sequences must agree with the pinned counts (length = nodes + leaves), start
at the root, only visit children of nodes visited before, and contain the
hit leaf.  The step aggregation is pinned on a hand-computed tile and by
conservation laws on real tiles."""
import numpy as np
import pytest

import oracle
from tests import handbvh
from tests.test_oracle_traversal import MINUS_X, PLUS_X, fray, orc

LEAF = 1 << 31


def seq(o, ray, mode=oracle.ORC_CLOSEST):
    ids, lens = o.visit_sequences(ray, mode)
    return [int(x) for x in ids[0, :lens[0]]]


def test_known_sequences_quirk_tree():
    # c=3 < M=4, +x: the order row [0,1,2,pad 0] is read backwards, so the
    # padding pushes child 0 (FastNode 1) first and child 1 (FastNode 2)
    # last: FastNode 2 is visited first, FastNode 1 (the tie winner) last
    assert seq(orc(handbvh.quirk_tree(False)), PLUS_X) == [0, 2 | LEAF, 1 | LEAF]
    # c=4: [0,1,2,3] backwards pushes children 1 (FastNode 2) then 0
    # (FastNode 1): FastNode 1 first, FastNode 2 last
    assert seq(orc(handbvh.quirk_tree(True)), PLUS_X) == [0, 1 | LEAF, 2 | LEAF]
    # -x reads the row forwards: push FastNode 1, then 2; pop 2 first
    assert seq(orc(handbvh.quirk_tree(True)), MINUS_X) == [0, 2 | LEAF, 1 | LEAF]


def test_known_sequences_stale_tree():
    o = orc(handbvh.stale_tree())
    origin = fray([0, 0, 0], [8, 0, 0])
    # root pushes A (FastNode 1); A's x row [0,2,3,1] read backwards pushes
    # slot 1 (FastNode 5), slot 3 (7, stale distance), slot 0 (4): pops 4, 7, 5
    assert seq(o, origin) == [0, 1, 4 | LEAF, 7 | LEAF, 5 | LEAF]
    # any-hit: slot order over all M slots pushes A (1) and the zero box slot
    # 3 = FastNode 4; 4 is popped first and occludes
    assert seq(o, origin, oracle.ORC_ANY) == [0, 4 | LEAF]
    # save distance: stale slot 3 carries t0 = 0, the x=3 leaf (t0 0.375 >
    # the hit's 0.25) is popped but skipped, so it is not a visit
    assert seq(o, origin, oracle.ORC_SAVE_DIST) == [0, 1, 4 | LEAF, 7 | LEAF]


def test_step_records_hand_tile():
    # ray 0: node A, leaf B; ray 1: node A, node C, leaf D; ray 2 not started
    A, B, C, D = 0, 5 | LEAF, 3, 9 | LEAF
    ids = np.array([[A, B, 0], [A, C, D], [7, 7, 7]], np.uint32)
    lens = np.array([2, 3, 3], np.uint32)
    (rec,) = oracle.tile_step_records(ids, lens, 3, [True, True, False])
    got = [tuple(int(v) for v in r) for r in rec]
    assert got == [
        (2, 1, 0, 0, 1),  # before the loop: 2 started, 1 not
        (2, 1, 0, 0, 0),  # both at A: one distinct node
        (1, 1, 1, 1, 1),  # ray 0 at leaf B (and done), ray 1 at node C
        (0, 0, 1, 1, 1),  # ray 1 at leaf D, done
    ]
    # nothing started: the single record before the loop
    (rec,) = oracle.tile_step_records(ids, lens, 3, [False] * 3)
    assert [tuple(int(v) for v in r) for r in rec] == [(0, 1, 0, 0, 3)]


@pytest.fixture(scope="module")
def scene(rt):
    tris = rt.interior_scene(5, 3000, 6)
    bvh = rt.Bvh(tris, 4, 4)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 48, 16)
    return bvh, rays


@pytest.mark.parametrize("mode", [oracle.ORC_CLOSEST, oracle.ORC_ANY, oracle.ORC_SAVE_DIST])
def test_sequences_agree_with_counts_and_tree(scene, mode):
    bvh, rays = scene
    o = oracle.Oracle.from_bvh(bvh)
    ids, lens = o.visit_sequences(rays, mode)
    r = o.trace(rays, mode=mode)
    c = r["counts"]
    assert np.array_equal(lens, c["nodes"] + c["leaves"])
    base, M = bvh.base, bvh.node_memory
    for i in range(0, len(rays), 7):
        s = ids[i, :lens[i]]
        assert s[0] == 0
        nodes = s[(s >> 31) == 0]
        assert len(nodes) == c["nodes"][i] and len(s) - len(nodes) == c["leaves"][i]
        seen = {0}
        for x in s[1:]:
            fid = int(x & 0x7FFFFFFF)
            parents = [p for p in seen if base[p] <= fid < base[p] + M]
            assert parents, (i, fid)
            if not x >> 31:
                seen.add(fid)
        if mode != oracle.ORC_ANY and r["hits"]["slot"][i] >= 0:
            leaf_ids = s[(s >> 31) == 1] & 0x7FFFFFFF
            assert r["hits"]["leaf"][i] in set(int(base[f]) for f in leaf_ids)
        if mode == oracle.ORC_ANY and r["occluded"][i]:
            assert s[-1] >> 31 == 1


def test_step_records_conservation(scene):
    bvh, rays = scene
    o = oracle.Oracle.from_bvh(bvh)
    ids, lens = o.visit_sequences(rays)
    recs = oracle.tile_step_records(ids, lens, 256, np.ones(len(rays), bool))
    assert len(recs) == len(rays) // 256
    for i, rec in enumerate(recs):
        v = lens[i * 256:(i + 1) * 256].astype(np.int64)
        assert len(rec) == v.max() + 1
        assert rec["terminations"].sum() == 256
        t = np.arange(len(rec))
        assert (t * rec["terminations"]).sum() == v.sum()
        assert np.array_equal((rec["node_work"] + rec["leaf_work"])[1:], [(v >= k).sum() for k in t[1:]])
        assert np.all(rec["node_unique"][1:] <= rec["node_work"][1:])
        assert np.all(rec["leaf_unique"] <= rec["leaf_work"])
        assert np.all((rec["node_unique"][1:] > 0) == (rec["node_work"][1:] > 0))
        assert rec["node_unique"][1] == 1  # every ray starts at the root
