"""Hand-built FastNode arrays (the layout of fastNodeManager.h:29-90,
1495-1673) for known-answer traversal tests.  Order rows are given
explicitly, as worked out by hand in each test."""
import math

import numpy as np


def roundup(x, g):
    return int(math.ceil(x / g) * g)


def tri_box(t):
    t = np.asarray(t, np.float32).reshape(3, 3)
    return t.min(0), t.max(0)


def build(M, Lm, G, spec):
    """spec: list of nodes in FastNode order.  Inner node:
    {"children": [ids...], "boxes": [(min3, max3), ...], "order": [[...],[...],[...]]}
    (order rows of length childCount; padded with 0 to M here).
    Leaf: {"tris": [9 floats, ...], "ids": [...]}.  Children of a node must be
    a contiguous id block (customTreeOrder)."""
    n = len(spec)
    bounds = np.zeros((n, 6 * M), np.float32)
    base = np.zeros(n, np.uint32)
    order = np.zeros((n, 3, M), np.int8)
    ctype = np.zeros(n, np.uint16)
    tri, ids = [], []
    for i, nd in enumerate(spec):
        if "children" in nd:
            ch = nd["children"]
            c = len(ch)
            assert ch == list(range(ch[0], ch[0] + c))
            nrs = roundup(c, G)
            for j, (bmin, bmax) in enumerate(nd["boxes"]):
                for a in range(3):
                    bounds[i, a * nrs + j] = bmin[a]
                    bounds[i, 3 * nrs + a * nrs + j] = bmax[a]
            if nrs != M:
                bounds[i, 6 * M - 2] = nrs
                bounds[i, 6 * M - 1] = np.nan
            base[i] = ch[0]
            for a in range(3):
                order[i, a, :c] = nd["order"][a]
            for j, cid in enumerate(ch):
                if "children" in spec[cid]:
                    ctype[i] |= 1 << j
        else:
            ts = np.asarray(nd["tris"], np.float32).reshape(-1, 9)
            pc = len(ts)
            lrs = roundup(pc, G)
            blk = np.zeros(9 * Lm, np.float32)
            for v in range(9):
                blk[v * lrs: v * lrs + pc] = ts[:, v]
            if lrs != Lm:
                blk[9 * Lm - 2] = lrs
                blk[9 * Lm - 1] = np.nan
            base[i] = len(tri) * 9 * Lm
            tri.append(blk)
            bounds[i, 0] = np.nan
            ids.extend(list(nd.get("ids", range(pc))) + [-1] * (Lm - pc))
    return dict(node_memory=M, leaf_memory=Lm, gang=G, bounds=bounds, base=base, order=order,
                child_type=ctype, tri_points=np.concatenate(tri), tri_ids=np.array(ids, np.int32))


def quirk_tree(full):
    """Root with three (full=False) or four (full=True) leaf children, M=4.
    Slots 0 and 1 hold the same triangle T in the plane x=5; slots 2 (and 3)
    hold triangles far off the ray.  All order rows are [0, 1, 2(, 3)]."""
    T = [5, 0, 0, 5, 1, 0, 5, 0, 1]
    far2 = [9, 5, 5, 9, 6, 5, 9, 5, 6]
    far3 = [9.5, 7, 7, 9.5, 8, 7, 9.5, 7, 8]
    leaves = [[T], [T], [far2]] + ([[far3]] if full else [])
    c = len(leaves)
    spec = [{"children": list(range(1, c + 1)), "boxes": [tri_box(l[0]) for l in leaves],
             "order": [list(range(c))] * 3}]
    spec += [{"tris": l, "ids": [10 + i]} for i, l in enumerate(leaves)]
    return build(4, 4, 4, spec)


def anyhit_tree(second_hit):
    """Root (M=4, c=3): slot 0 = leaf at x=7, slot 1 = leaf with T at x=5,
    slot 2 = far leaf.  The slot-0 triangle is hit only if second_hit.
    Order rows: x [1, 0, 2] (centre 10.5 < 14.5 < 18.5), y and z [0, 1, 2]."""
    T5 = [5, 0, 0, 5, 1, 0, 5, 0, 1]
    T7 = [7, 0, 0, 7, 1, 0, 7, 0, 1] if second_hit else [7, 1, 1, 7, 0, 1, 7, 1, 0]
    far = [9, 5, 5, 9, 6, 5, 9, 5, 6]
    leaves = [[T7], [T5], [far]]
    spec = [{"children": [1, 2, 3], "boxes": [tri_box(l[0]) for l in leaves],
             "order": [[1, 0, 2], [0, 1, 2], [0, 1, 2]]}]
    spec += [{"tris": l, "ids": [20 + i]} for i, l in enumerate(leaves)]
    return build(4, 4, 4, spec)


def stale_tree():
    """Root (c=3, M=4: slot 3 is a zero box) -> inner node A (c=4) -> 4 leaves.
    A ray from the origin along +x hits the root's zero box (stale distance
    at slot 3) and A's slots 0 and 1 only; A's slot 3 is pushed anyway."""
    def sq(x, y0):
        return [x, y0 - 1, y0 - 1, x, y0 + 2, y0 - 1, x, y0 - 1, y0 + 2]
    a_leaves = [[sq(2, 0)], [sq(3, 0)], [sq(2, 10)], [sq(2, 20)]]
    far1, far2 = [sq(1, 30)], [sq(1, 40)]
    a_box = (np.float32([1, -1, -1]), np.float32([4, 2, 2]))
    spec = [
        {"children": [1, 2, 3], "boxes": [a_box, tri_box(far1[0]), tri_box(far2[0])],
         "order": [[0, 1, 2], [0, 1, 2], [0, 1, 2]]},
        {"children": [4, 5, 6, 7], "boxes": [tri_box(l[0]) for l in a_leaves],
         "order": [[0, 2, 3, 1], [0, 1, 2, 3], [0, 1, 2, 3]]},
        {"tris": far1, "ids": [1]},
        {"tris": far2, "ids": [2]},
    ] + [{"tris": l, "ids": [30 + i]} for i, l in enumerate(a_leaves)]
    return build(4, 4, 4, spec)


def marker_tree():
    """M = 8, G = 4, Lm = 8: loop counts come from the NaN markers.
    Node 0 (root, 3 children -> nodeRealSize 4 != M, marker, k = 4):
      slot 0 = node 1, inner, box x [3, 8], y/z [0, 4]
      slot 1 = node 2, leaf A, 3 triangles (leafRealSize 4, marker, k = 4)
               in the plane y = 5.5 region: box x [1, 2], y/z [0, 6]; the
               triangles lie at y, z in [5, 6], off every test ray
      slot 2 = node 3, leaf B, 5 copies of one triangle at x = 9.5 covering
               y + z <= 4 (leafRealSize 8 == Lm, no marker, k = 8)
      order x [1, 0, 2] (keys 3.5, 11.5, 19.5); y, z [0, 2, 1] (4.5, 4.5 tie
      broken by index, then 6.5)
    Node 1 (6 children -> nodeRealSize 8 == M, no marker, k = 8):
      slots 0..4 = nodes 4..8, leaves of one triangle at x = 3, 4, 5, 6, 7
               covering y + z <= 1 (leafRealSize 4, marker, k = 4)
      slot 5 = node 9, leaf of 6 triangles at x = 7.5 (y + z <= 1,
               leafRealSize 8, no marker, k = 8)
      order rows [0, 1, 2, 3, 4, 5] on every axis (y/z keys all tie)
    Leaf blocks in FastNode order: A = 0, B = 1, nodes 4..9 = 2..7, so a
    leaf's primIdBegin is 72 * block."""
    def wall(x, size):
        return [x, 0, 0, x, size, 0, x, 0, size]

    def box(x0, x1, hi):
        return (np.float32([x0, 0, 0]), np.float32([x1, hi, hi]))

    a_tris = [[1.0, 5, 5, 1.5, 6, 5, 2.0, 5, 6], [1.5, 5, 5, 1.5, 6, 5, 1.5, 5, 6], [2.0, 5, 5, 1.0, 6, 5, 1.0, 5, 6]]
    kids = [[wall(float(x), 1)] for x in (3, 4, 5, 6, 7)] + [[wall(7.5, 1)] * 6]
    spec = [
        {"children": [1, 2, 3], "boxes": [box(3, 8, 4), box(1, 2, 6), box(9.5, 9.5, 4)],
         "order": [[1, 0, 2], [0, 2, 1], [0, 2, 1]]},
        {"children": [4, 5, 6, 7, 8, 9], "boxes": [box(3, 3, 1), box(4, 4, 1), box(5, 5, 1), box(6, 6, 1),
                                                   box(7, 7, 1), box(7.5, 7.5, 1)],
         "order": [[0, 1, 2, 3, 4, 5]] * 3},
        {"tris": a_tris, "ids": [100, 101, 102]},
        {"tris": [wall(9.5, 4)] * 5, "ids": [110, 111, 112, 113, 114]},
    ] + [{"tris": t, "ids": [120 + 10 * i + j for j in range(len(t))]} for i, t in enumerate(kids)]
    return build(8, 8, 4, spec)
