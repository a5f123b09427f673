"""Host entry points of the reference, driving the GPU traversal.

  read_config        RayTracer::readConfig   (src/rayTracer.cpp:756-877, src/configUtil.h:6-58)
  RayTracer.run      the N/L sweep            (rayTracer.cpp:97-122 clamps, :568-576 loop)
  RayTracer.render   RayTracer::renderImage   (rayTracer.cpp:588-745, performance branch)
  CameraFast         CameraFast::renderImages / renderImage (src/cameraFast.h:49-582)

Scenes are procedural (the reference loads glTF models that are not part of
the repository); everything else keeps the reference's semantics: substring
key matching, `#` comments, the 1 warm-up + 5 timed renders median protocol,
and the `_Perf.txt` fields.  Traversal runs on the GPU through librt.so.
"""
import math
import os
import time

import numpy as np

from .host import (Bvh, Scene, camera_rays, interior_scene, load_sah_factor, make_rays,
                   spawn_shadow, spheres_scene)
from ._lib import RAY_DTYPE

INT_KEYS = ["minLeafSize", "maxLeafSize", "leafStep", "minBranch", "maxBranch", "branchStep",
            "subdivisionStart", "subdivisionEnd", "subdivisionStep", "renderType",
            "ambientSampleCount", "workGroupSize", "leafSplitOption", "xRes", "yRes"]
BOOL_KEYS = ["saveImage", "saveDepthDetailedImage", "bvhAnalysis", "saveBvhImage", "sortEachSplit",
             "renderAnalysisImage", "castShadows", "doPerformanceTest", "doNodeMemoryTest",
             "doLeafMemoryTest", "saveRayTimes", "doWorkGroupAnalysis", "saveDistance", "wideRender",
             "wideAlternative", "renderAllOptions", "doCacheAnalysis"]


def _stoi(s):
    """std::stoi: skip leading whitespace, read an optional sign and digits."""
    s = s.lstrip()
    j = 0
    if j < len(s) and s[j] in "+-":
        j += 1
    while j < len(s) and s[j].isdigit():
        j += 1
    if j == 0 or s[:j] in "+-":
        raise ValueError(f"stoi: no conversion for {s!r}")
    return int(s[:j])


def read_config(path="config.txt", defaults=None):
    """RayTracer::readConfig: for every non-empty line that does not start with
    '#', every key that occurs anywhere in the line (substring match) takes the
    value after the first '='.  Booleans look for "true" / "false" anywhere in
    the line; a line with neither sets false (with a warning)."""
    cfg = dict(defaults or {})
    cfg.setdefault("scenarios", [])
    cfg.setdefault("cacheSizes", [])
    warnings = []
    with open(path) as f:
        for line in f.read().splitlines():
            if not line or line[0] == "#":
                continue
            for k in INT_KEYS:
                if k in line:
                    cfg[k] = _stoi(line[line.find("=") + 1:])
            for k in BOOL_KEYS:
                if k in line:
                    if "true" in line:
                        cfg[k] = True
                    elif "false" in line:
                        cfg[k] = False
                    else:
                        warnings.append(f"{k}value written wrong -> default = false")
                        cfg[k] = False
            if "scenario" in line:
                s = line[line.find("=") + 1:]
                cfg["scenarios"] += [_stoi(t) for t in s.split(",")]
            if cfg.get("doCacheAnalysis") and "cacheSize" in line:
                for t in line[line.find("=") + 1:].split(","):
                    v = _stoi(t)
                    if v % 8 == 0:
                        cfg["cacheSizes"].append(v)
                    else:
                        warnings.append(f"cachesize must be divideable by 8: {v} is ignored")
                if not cfg["cacheSizes"]:
                    cfg["cacheSizes"] = [256]
    cfg["_warnings"] = warnings
    return cfg


def sanity_clamp(cfg):
    """rayTracer.cpp:104-122 settings sanity checks."""
    cfg["maxBranch"] = max(cfg["maxBranch"], cfg["minBranch"])
    cfg["minBranch"] = min(cfg["maxBranch"], cfg["minBranch"])
    cfg["maxLeafSize"] = max(cfg["maxLeafSize"], cfg["minLeafSize"])
    cfg["minLeafSize"] = min(cfg["maxLeafSize"], cfg["minLeafSize"])
    return cfg


def sweep(cfg):
    """The (leafSize, branchingFactor) loop of rayTracer.cpp:568-576."""
    for leaf in range(cfg["minLeafSize"], cfg["maxLeafSize"] + 1, cfg["leafStep"]):
        for branch in range(cfg["minBranch"], cfg["maxBranch"] + 1, cfg["branchStep"]):
            yield leaf, branch


def _sphere_dirs(index, sample, normal):
    """Deterministic uniform directions flipped into the normal's hemisphere
    (the sampler shape of Camera::sampleHemisphere, camera.cpp:36-52, fed by
    a PCG-style hash instead of the platform std::hash)."""
    h = (index.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
         + np.uint64(sample) * np.uint64(0xBF58476D1CE4E5B9) + np.uint64(1))
    h ^= h >> np.uint64(31)
    h *= np.uint64(0x94D049BB133111EB)
    h ^= h >> np.uint64(29)
    u = ((h >> np.uint64(40)).astype(np.float64)) / float(1 << 24)
    v = (((h >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64)) / float(1 << 24)
    phi = np.arccos(1 - 2 * u)
    theta = 2 * math.pi * v
    d = np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], 1)
    flip = (d * normal).sum(1) <= 0
    d[flip] = -d[flip]
    return d.astype(np.float32)


class CameraFast:
    """CameraFast::renderImages: for every camera one warm-up render, then
    `samples` timed renders; the median (by total time) is kept.  A render =
    primary closest-hit rays for the whole image (one rt_trace) + hit
    reconstruction + ambientSampleCount any-hit rays per hit of length
    ambientDistance (one rt_trace per sample)."""

    def __init__(self, positions, targets, width, height, workgroup=16, focal=0.866):
        self.positions, self.targets = positions, targets
        self.width, self.height, self.wg, self.focal = width, height, workgroup, focal

    def render_image(self, scene, bvh, cam, ambient_samples, ambient_distance, mode="closest"):
        t0 = time.perf_counter()
        rays = camera_rays(self.positions[cam], self.targets[cam], self.width, self.height, self.wg,
                           focal=self.focal)
        hits = scene.intersect(rays, mode=mode)
        t_primary = time.perf_counter() - t0
        occl = np.zeros(len(rays), np.uint16)
        hit = hits["slot"] >= 0
        if ambient_samples and hit.any():
            nrm, pos = bvh.surface(rays, hits)
            idx = np.nonzero(hit)[0]
            origin = pos[idx] + nrm[idx] * np.float32(0.001)
            for a in range(ambient_samples):
                d = _sphere_dirs(idx, a, nrm[idx])
                amb = make_rays(origin, origin + d, tmax=ambient_distance)
                occl[idx] += scene.intersect(amb, mode="any")
        total = time.perf_counter() - t0
        image = ((1 - occl / max(1, ambient_samples)) * 255).astype(np.uint8) if ambient_samples else None
        return total, t_primary, hits, image

    def render_images(self, scene, bvh, ambient_samples=1, ambient_distance=1.0, samples=5,
                      mode="closest"):
        totals, primaries = [], []
        for cam in range(len(self.positions)):
            self.render_image(scene, bvh, cam, ambient_samples, ambient_distance, mode)  # warm-up
            res = sorted(self.render_image(scene, bvh, cam, ambient_samples, ambient_distance, mode)[:2]
                         for _ in range(samples))
            med = res[samples // 2]
            totals.append(med[0])
            primaries.append(med[1])
        return float(np.mean(totals)), float(np.mean(primaries))


def write_perf(path, name, problem, gang, node_memory, leaf_memory, branching, leaf_size,
               total, ray_sum, tri_sum, avg_depth, tri_count):
    """The _Perf.txt fields of cameraFast.h:147-161."""
    with open(path, "w") as f:
        f.write(f"scenario {name} with branching factor of {branching} and leafsize of {leaf_size}\n")
        f.write(f"with a gang size of {gang}, a node meory of {node_memory}, and a leaf memory of {leaf_memory}\n")
        f.write(f"Raytracer total time: {total}\n")
        f.write(f"Time for all rays (SUM): {ray_sum}\n")
        f.write(f"Time for triangle intersections (SUM): {tri_sum}\n")
        f.write(f"Time all rays(sum) - triangle(sum): {ray_sum - tri_sum}\n")
        f.write(f"Average BVH depth: {avg_depth}\n")
        f.write(f"Triangle Count: {tri_count}\n")


def _info_block(f, label, st):
    rays = max(1, st["rays"])
    f.write(f"{label} intersections node: {st['node_visits'] / rays}\n")
    f.write(f"{label} aabb intersections: {st['aabb_tests'] / rays}\n")
    f.write(f"{label} aabb success ration: {st['aabb_hits'] / max(1, st['aabb_tests'])}\n")
    waste = (st["leaf_visits"] + st["node_visits"] - st["rays"]) / max(1, st["node_visits"] * st["branching"])
    f.write(f"{label} waste factor: {1 - waste}\n")
    f.write(f"{label} intersections leaf: {st['leaf_visits'] / rays}\n")
    f.write(f"{label} primitive intersections: {st['tri_tests'] / rays}\n")
    f.write(f"{label} primitive success ratio: {st['tri_hits'] / max(1, st['tri_tests'])}\n")


def _trim(h):
    nz = np.nonzero(h)[0]
    return h[: nz[-1] + 1] if len(nz) else h[:1]


def write_info(path, name, branching, leaf_size, primary, secondary):
    """The _Info.txt of the reference's instrumented renderer
    (cameraData.h:241-329), from rt_trace_histograms counters of a primary
    and a secondary (shadow) batch: per-ray averages, success ratios, the
    waste factor, child/leaf fullness and visits per depth, with the same
    line keys, so the reference's Analysis/DataManager.py reads it."""
    p = dict(primary, branching=branching)
    s = dict(secondary, branching=branching)
    with open(path, "w") as f:
        f.write(f"scenario {name} with branching factor of {branching} and leafsize of {leaf_size}\n")
        f.write("intersections counts are normalized per Ray\n\n")
        _info_block(f, "primary", p)
        f.write("\n")
        _info_block(f, "secondary", s)
        f.write("\n")
        f.write("intersections with nodes with x children :\n")
        cf = _trim(p["child_fullness"])
        f.write(f"average child fullness: {float((cf * np.arange(len(cf))).sum() / max(1, cf.sum())):.6f}\n")
        for i, v in enumerate(cf):
            f.write(f"{i} : {v / max(1, p['rays'])}\n")
        f.write("\n")
        f.write("intersections with leaf nodes with x primitives :\n")
        lf = _trim(p["leaf_fullness"])
        f.write(f"average leaf fullness: {float((lf * np.arange(len(lf))).sum() / max(1, lf.sum())):.6f}\n")
        for i, v in enumerate(lf):
            f.write(f"{i} : {v / max(1, p['rays'])}\n")
        for label, st, key in (("primary node", p, "node_depth"), ("primary leaf", p, "leaf_depth"),
                               ("secondary node", s, "node_depth"), ("secondary leaf", s, "leaf_depth")):
            f.write("\n")
            f.write(f"{label} intersections at depth x :\n")
            for i, v in enumerate(_trim(st[key])):
                f.write(f"{i} : {v / max(1, st['rays'])}\n")


SCENES = {
    "spheres": dict(make=lambda: spheres_scene(0), cam=((0.0, 4.0, 32.0), (0.0, 0.0, 0.0))),
    "interior": dict(make=lambda: interior_scene(2, 1_000_000), cam=((0.0, 5.0, -14.0), (0.0, 3.0, 0.0))),
}


class RayTracer:
    """RayTracer::run with doPerformanceTest: for every (L, N) of the sweep,
    load the SAH factor, build the BVH, flatten, upload and render."""

    def __init__(self, cfg, gang=4, out_dir="perf"):
        self.cfg = sanity_clamp(dict(cfg))
        self.gang = gang
        self.out_dir = out_dir

    def render(self, name, tris, branching, leaf_size, positions, targets):
        cfg = self.cfg
        sah = load_sah_factor(leaf_size, branching, self.gang)
        bvh = Bvh(tris, branching, leaf_size, gang=self.gang, sort_each_split=cfg.get("sortEachSplit", True),
                  leaf_split_option=cfg.get("leafSplitOption", 1), sah_factor=sah)
        scene = Scene(bvh)
        # ambientDistance = cbrt(root volume) / 10 (rayTracer.cpp:607)
        lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
        ambient = float(np.cbrt(np.prod((hi - lo).astype(np.float64))) / 10.0)
        cam = CameraFast(positions, targets, cfg.get("xRes", 1920), cfg.get("yRes", 1088),
                         cfg.get("workGroupSize", 16))
        mode = "save_distance" if cfg.get("saveDistance") else "closest"
        total, primary = cam.render_images(scene, bvh, cfg.get("ambientSampleCount", 1), ambient, mode=mode)
        os.makedirs(self.out_dir, exist_ok=True)
        problem = f"_b{branching}_l{leaf_size}"
        prefix = f"_mb{bvh.node_memory}_ml{bvh.leaf_memory}"
        write_perf(os.path.join(self.out_dir, f"{name}{problem}{prefix}_Perf.txt"), name, problem, self.gang,
                   bvh.node_memory, bvh.leaf_memory, branching, leaf_size, total, primary, 0.0,
                   bvh.average_depth, bvh.triangle_count)
        scene.close()
        return total

    def run(self, scenes=("spheres",)):
        results = {}
        for name in scenes:
            sc = SCENES[name]
            tris = sc["make"]()
            pos, tgt = sc["cam"]
            for leaf, branch in sweep(self.cfg):
                results[(name, branch, leaf)] = self.render(name, tris, branch, leaf, [pos], [tgt])
        return results
