"""Writes config1_n4l4.npz: the regression fixture of the 256x256 config-1
frame (spheres scene seed 0, SAH N4L4 gang 4, camera (0,4,32) -> (0,0,0)).

The reference ships no golden vectors for this path (its only test pins a
cache simulator), so this fixture is this repository's own: it freezes the
host builder's topology (node count, depth) and the oracle's per-ray results
for every 16th ray, plus SHA-256 digests of the full result arrays.  A change
to the builder, the ray generation or the oracle shows up as a mismatch.

Run from the repository root: python tests/golden/make_config1.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def frame():
    import bvhraytracer_amd as rt
    import oracle

    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 4, 4)
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 256, 256, 16)
    o = oracle.Oracle.from_bvh(bvh)
    closest = o.trace(rays)
    return tris, bvh, rays, closest


def digests(tris, bvh, rays, closest):
    def h(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

    return {"tris": h(tris), "bounds": h(bvh.bounds), "tri_points": h(bvh.tri_points), "rays": h(rays),
            "hits": h(closest["hits"]), "counts": h(closest["counts"])}


def main():
    tris, bvh, rays, closest = frame()
    sel = np.arange(0, len(rays), 16)
    d = digests(tris, bvh, rays, closest)
    np.savez_compressed(os.path.join(HERE, "config1_n4l4.npz"),
                        node_count=np.int64(bvh.node_count), max_depth=np.int64(bvh.max_depth),
                        ray_index=sel, rays=rays[sel], hits=closest["hits"][sel], counts=closest["counts"][sel],
                        digests=np.array([f"{k}={v}" for k, v in sorted(d.items())]))
    print("wrote", os.path.join(HERE, "config1_n4l4.npz"), bvh.node_count, bvh.max_depth)


if __name__ == "__main__":
    main()
