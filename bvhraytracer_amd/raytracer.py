"""Host entry points of the reference, driving the GPU traversal.

  read_config        RayTracer::readConfig   (src/rayTracer.cpp:756-877, src/configUtil.h:6-58)
  RayTracer.run      the N/L sweep            (rayTracer.cpp:97-122 clamps, :568-576 loop)
  RayTracer.render   RayTracer::renderImage   (rayTracer.cpp:588-745, performance branch)
  CameraFast         CameraFast::renderImages / renderImage (src/cameraFast.h:49-582)
  write_png          the _Perf.png of cameraFast.h:829-852 (tile order -> rows, RGBA PNG)
  workgroup_analysis the wide render's per-step tile statistics and the four
                     WorkGroup*.txt files of cameraData.h:458-560,660-1109
  analysis_render    CameraData's analysis render: _Info.txt (+ workgroup files)

Scenes are procedural (the reference loads glTF models that are not part of
the repository); everything else keeps the reference's semantics: substring
key matching, `#` comments, the 1 warm-up + 5 timed renders median protocol,
and the `_Perf.txt` fields.  Traversal runs on the GPU through librt.so.
"""
import os
import struct
import time
import zlib

import numpy as np

from .host import (Bvh, Scene, camera_rays, interior_scene, load_sah_factor, spawn_ambient, spheres_scene)
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


def tile_order_to_rows(values, width, height, wg):
    """Per-pixel values in CameraFast tile order (the order of camera_rays:
    wg x wg tiles left to right, tile rows top to bottom, lanes row-major in a
    tile) to a (height, width, ...) row-major image: the reordering before
    cameraFast.h:833-849 saves the image, extended to a partial last tile
    row (whose tiles hold height % wg rows)."""
    values = np.asarray(values)
    if width % wg:
        raise ValueError("width must be a multiple of the workgroup size")
    if len(values) != width * height:
        raise ValueError(f"expected {width * height} pixels, got {len(values)}")
    tiles_x, full_rows = width // wg, height // wg
    h, w = np.meshgrid(np.arange(height, dtype=np.int64), np.arange(width, dtype=np.int64), indexing="ij")
    lane = (h % wg) * wg + w % wg
    k = ((h // wg) * tiles_x + w // wg) * wg * wg + lane
    part = h >= full_rows * wg
    if part.any():
        rows_left = height - full_rows * wg
        k[part] = (full_rows * tiles_x * wg * wg + (w[part] // wg) * rows_left * wg
                   + (h[part] % wg) * wg + w[part] % wg)
    return values[k]


def _png_chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path, rgba):
    """An 8-bit RGBA PNG of a (height, width, 4) uint8 image (lodepng::encode
    of the reference's encodeTwoSteps, util.h:52-61, takes the same RGBA
    buffer; the pixels decode identically, the compressed bytes differ)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("expected a (height, width, 4) image")
    height, width = rgba.shape[:2]
    raw = b"".join(b"\x00" + rgba[r].tobytes() for r in range(height))  # filter type 0 per row
    png = (b"\x89PNG\r\n\x1a\n"
           + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0))
           + _png_chunk(b"IDAT", zlib.compress(raw, 6))
           + _png_chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)


def ambient_distance(tris):
    """ambientDistance of RayTracer::renderImage (rayTracer.cpp:607): cbrt of
    the root box volume (Aabb::getVolume, aabb.h:136-140) / 10, in float32,
    the root box being the triangles' bounds."""
    pts = np.asarray(tris, np.float32).reshape(-1, 3)
    d = pts.max(0) - pts.min(0)
    vol = np.float32(np.float32(d[0] * d[1]) * d[2])
    return float(np.float32(np.cbrt(vol) / np.float32(10)))


def ambient_image(occluded, hit, samples):
    """Gray value per pixel (cameraFast.h:404-439): a hit pixel is
    (uint8)((1 - occluded / samples) * 255) in float32, a miss is 0.  With
    no samples the reference converts NaN, which gives 0 on x86."""
    img = np.zeros(len(hit), np.uint8)
    if samples:
        factor = np.float32(1) - occluded.astype(np.float32) / np.float32(samples)
        img[hit] = (factor[hit] * np.float32(255)).astype(np.uint8)
    return img


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
            for a in range(ambient_samples):
                amb, src = spawn_ambient(nrm, pos, hits, a, ambient_distance)
                occl[src] += scene.intersect(amb, mode="any")
        total = time.perf_counter() - t0
        return total, t_primary, hits, ambient_image(occl, hit, ambient_samples)

    def render_image_device(self, scene, cam, ambient_samples, ambient_distance, mode="closest", wide=False):
        """The same frame as render_image with every step on the GPU
        (rt_render_ambient; identical gray values).  Times are the GPU's
        (HIP events): (total s, primary s, None, image)."""
        gray, (p_ms, t_ms) = scene.render_ambient(self.positions[cam], self.targets[cam], self.width, self.height,
                                                  self.wg, ambient_samples, ambient_distance, mode=mode, wide=wide,
                                                  focal=self.focal)
        return t_ms / 1000.0, p_ms / 1000.0, None, gray

    def save_image(self, path, name, cam, image):
        """<path>/<name>_c<cam>_Perf.png (cameraFast.h:829-852), gray RGBA."""
        g = tile_order_to_rows(image, self.width, self.height, self.wg)
        rgba = np.empty(g.shape + (4,), np.uint8)
        rgba[..., :3] = g[..., None]
        rgba[..., 3] = 255
        out = os.path.join(path, f"{name}_c{cam}_Perf.png")
        write_png(out, rgba)
        return out

    def render_images(self, scene, bvh, ambient_samples=1, ambient_distance=1.0, samples=5,
                      mode="closest", save_image=False, path=".", name="image", device=True):
        """With save_image, one untimed render per camera first writes its
        _Perf.png (cameraFast.h:81-89).  device: every frame on the GPU
        (render_image_device, GPU times); else the host-driven render_image
        (wall times)."""
        def one(cam):
            if device:
                return self.render_image_device(scene, cam, ambient_samples, ambient_distance, mode)
            return self.render_image(scene, bvh, cam, ambient_samples, ambient_distance, mode)

        totals, primaries = [], []
        if save_image:
            os.makedirs(path, exist_ok=True)
            for cam in range(len(self.positions)):
                self.save_image(path, name, cam, one(cam)[3])
        for cam in range(len(self.positions)):
            one(cam)  # warm-up
            res = sorted(one(cam)[:2] for _ in range(samples))
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
        # the reference streams floats: %g with 6 significant digits
        total, ray_sum, tri_sum = np.float32(total), np.float32(ray_sum), np.float32(tri_sum)
        f.write(f"Raytracer total time: {_g(total)}\n")
        f.write(f"Time for all rays (SUM): {_g(ray_sum)}\n")
        f.write(f"Time for triangle intersections (SUM): {_g(tri_sum)}\n")
        f.write(f"Time all rays(sum) - triangle(sum): {_g(np.float32(ray_sum - tri_sum))}\n")
        f.write(f"Average BVH depth: {_g(np.float32(avg_depth))}\n")
        f.write(f"Triangle Count: {tri_count}\n")


def _info_block(f, label, st, branching):
    """One counter block of _Info.txt (cameraData.h:247-265), float32 as
    there: factor = 1 / (float)rays, uint64 counts converted to float."""
    f32 = np.float32
    factor = f32(1) / f32(st["rays"])

    def per_ray(v):
        return _g(f32(f32(int(v)) * factor))

    def ratio(a, b):
        return _g(f32(f32(int(a)) / f32(int(b))))

    f.write(f"{label} intersections node: {per_ray(st['node_visits'])}\n")
    f.write(f"{label} aabb intersections: {per_ray(st['aabb_tests'])}\n")
    f.write(f"{label} aabb success ration: {ratio(st['aabb_hits'], st['aabb_tests'])}\n")
    waste = f32(f32(int(st["leaf_visits"] + st["node_visits"] - st["rays"])) / f32(int(st["node_visits"]) * branching))
    f.write(f"{label} waste factor: {_g(f32(1) - waste)}\n")
    f.write(f"{label} intersections leaf: {per_ray(st['leaf_visits'])}\n")
    f.write(f"{label} primitive intersections: {per_ray(st['tri_tests'])}\n")
    f.write(f"{label} primitive success ratio: {ratio(st['tri_hits'], st['tri_tests'])}\n")


def _fullness(f, title, avg_key, counts, size, factor):
    """A fullness section (cameraData.h:270-292): the average is a float sum
    of count * i over the counts' int total, written with std::to_string."""
    f32 = np.float32
    c = [int(v) for v in counts[:size]] + [0] * max(0, size - len(counts))
    acc = f32(0)
    for i, v in enumerate(c):
        acc = f32(acc + f32(v * i))
    total = sum(c) & 0xFFFFFFFF  # std::accumulate(..., 0) sums into an int
    total = total - (1 << 32) if total >= 1 << 31 else total
    acc = f32(acc / f32(total))
    f.write(f"{title}\n")
    f.write("%s%f\n" % (avg_key, float(acc)))
    for i, v in enumerate(c):
        f.write(f"{i} : {_g(f32(f32(v) * factor))}\n")


def write_info(path, name, branching, leaf_size, primary, secondary, bvh_depth=None):
    """The _Info.txt of the reference's instrumented renderer
    (cameraData.h:241-329), from rt_trace_histograms counters of the primary
    rays and of the secondary (ambient) rays: per-ray averages, success
    ratios, the waste factor, child fullness 0..N and leaf fullness 0..L of
    the primary rays, and visits per depth 0..bvh_depth-1 (bvhDepth, the
    tree depth + 1), in float32 with the reference's line keys and number
    formats, so its Analysis/DataManager.py reads it.  Every ray, test and
    visit count must be nonzero (the reference divides by them)."""
    f32 = np.float32
    if bvh_depth is None:
        nz = np.nonzero(np.asarray(primary["node_depth"]) + np.asarray(primary["leaf_depth"]))[0]
        bvh_depth = int(nz[-1]) + 1 if len(nz) else 1
    with open(path, "w") as f:
        f.write(f"scenario {name} with branching factor of {branching} and leafsize of {leaf_size}\n")
        f.write("intersections counts are normalized per Ray\n\n")
        _info_block(f, "primary", primary, branching)
        f.write("\n")
        _info_block(f, "secondary", secondary, branching)
        f.write("\n")
        factor = f32(1) / f32(primary["rays"])
        _fullness(f, "intersections with nodes with x children :", "average child fullness: ",
                  primary["child_fullness"], branching + 1, factor)
        f.write("\n")
        _fullness(f, "intersections with leaf nodes with x primitives :", "average leaf fullness: ",
                  primary["leaf_fullness"], leaf_size + 1, factor)
        for label, st, key in (("primary node", primary, "node_depth"), ("primary leaf", primary, "leaf_depth"),
                               ("secondary node", secondary, "node_depth"),
                               ("secondary leaf", secondary, "leaf_depth")):
            fac = f32(1) / f32(st["rays"])
            f.write("\n")
            f.write(f"{label} intersections at depth x :\n")
            for i in range(bvh_depth):
                v = int(st[key][i]) if i < len(st[key]) else 0
                f.write(f"{i} : {_g(f32(f32(v) * fac))}\n")


def _g(x):
    """std::ostream << float / double with the default format (%g, 6
    significant digits)."""
    return "%g" % float(x)


def _whisker_rows(tiles, wgsq):
    """Per-tile (median, min, max, avg, sd) of the step at which each ray
    finished (cameraData.h:703-779): the depth list expands the
    terminations per step; avg and the variance are float32 as there (each
    square comes from std::pow in double and is added to the float sum, one
    rounding per ray, in ray order; vectorised over tiles)."""
    med_id = int(np.float32(wgsq) * np.float32(0.5))
    depth = np.zeros((len(tiles), wgsq), np.int64)
    work = np.zeros(len(tiles), np.int64)
    for i, rec in enumerate(tiles):
        term = rec["terminations"].astype(np.int64)
        dl = np.repeat(np.arange(len(term), dtype=np.int64), term)
        if len(dl) < wgsq:
            raise ValueError(f"tile {i}: {len(dl)} terminations, expected at least {wgsq}")
        depth[i] = dl[:wgsq]
        work[i] = int((np.arange(len(term)) * term).sum())
    avg = (work.astype(np.float32) / np.float32(wgsq)).astype(np.float32)
    var = np.zeros(len(tiles), np.float32)
    for j in range(wgsq):
        d = (depth[:, j].astype(np.float32) - avg).astype(np.float32).astype(np.float64)
        var = (var.astype(np.float64) + d * d).astype(np.float32)
    var = (var / np.float32(wgsq)).astype(np.float32)
    sd = np.sqrt(var).astype(np.float32)
    rows = [(int(depth[i, med_id]), int(depth[i, 0]), int(depth[i, wgsq - 1]), avg[i], sd[i])
            for i in range(len(tiles))]
    # std::sort by (median, max); the reference leaves ties unordered, here
    # they keep the tile order
    return sorted(rows, key=lambda r: (r[0], r[2]))


def _step_averages(tiles, j, n_tiles_f, wgsq):
    """One row half of _WorkGroupData.txt for step j (cameraData.h:846-949):
    averages over the tiles still working at step j, unique counts over the
    tiles with a nonzero unique count, terminations over all tiles."""
    live = [t for t in tiles if len(t) > j]
    acc = {f: np.float32(0) for f in ("node_work", "leaf_work", "node_unique", "leaf_unique", "terminations")}
    for t in live:
        for f in acc:
            acc[f] = np.float32(acc[f] + np.float32(int(t[f][j])))
    un = [int(t["node_unique"][j]) for t in live if t["node_unique"][j] != 0]
    ul = [int(t["leaf_unique"][j]) for t in live if t["leaf_unique"][j] != 0]
    cnt = np.float32(max(len(live), 1))
    return (np.float32(acc["node_work"] / cnt), np.float32(acc["node_unique"] / np.float32(max(len(un), 1))),
            np.float32(acc["leaf_work"] / cnt), np.float32(acc["leaf_unique"] / np.float32(max(len(ul), 1))),
            np.float32(acc["terminations"] / n_tiles_f),
            max(un) if un else 0, min(un) if un else 0, max(ul) if ul else 0, min(ul) if ul else 0)


def write_workgroup_analysis(path, name, problem, camera_id, width, height, wg, primary, secondary,
                             primary_visits, secondary_visits, wide_alternative=True):
    """The workgroup analysis files of CameraData::workGroupAnalysisPerImage
    (cameraData.h:660-1109) under path/WorkGroupSize_<wg>_Version_<0|1>/:
    _PrimaryWorkGroupWiskerPlot.txt, _SecondaryWorkGroupWiskerPlot.txt,
    _WorkGroupData.txt and _WorkGroupUniqueWork.txt.

    primary / secondary: per full tile, the STEP_DTYPE records of
    Scene.workgroup_steps (secondary: the records of every ambient sample
    concatenated, as the reference appends them).  primary_visits /
    secondary_visits: (nodes, leaves) per ray in tile order (secondary summed
    over the samples).  Returns the four paths."""
    wgsq = wg * wg
    n_tiles = (width // wg) * (height // wg)
    if len(primary) != n_tiles or len(secondary) != n_tiles:
        raise ValueError(f"expected {n_tiles} tiles")
    d = os.path.join(path, f"WorkGroupSize_{wg}_Version_{int(bool(wide_alternative))}")
    os.makedirs(d, exist_ok=True)
    stem = os.path.join(d, f"{name}{problem}_c{camera_id}")
    out = [stem + "_PrimaryWorkGroupWiskerPlot.txt", stem + "_SecondaryWorkGroupWiskerPlot.txt",
           stem + "_WorkGroupData.txt", stem + "_WorkGroupUniqueWork.txt"]
    for fn, tiles in ((out[0], primary), (out[1], secondary)):
        with open(fn, "w") as f:
            f.write("median, min, max, lowerStdDeviation, upperStdDeviation\n")
            for med, mn, mx, avg, sd in _whisker_rows(tiles, wgsq):
                f.write(f"{med}, {mn}, {mx}, {_g(avg)}, {_g(sd)}\n")
    # the reference's size loop tests the primary size twice, so the steps
    # are those of the longest primary tile
    max_size = max(len(t) for t in primary)
    n_tiles_f = np.float32(np.float32(width // wg) * np.float32(height // wg))
    with open(out[2], "w") as f:
        f.write("stepId, avgPrimaryNodeWork, avgPrimaryNodeUnique, avgPrimaryLeafWork, avgPrimaryLeafUnique, "
                "avgPrimaryRayTermination, primaryNodeWorkMax, primaryNodeWorkMin, primaryLeafWorkMax, "
                "primaryLeafWorkMin, avgSecondaryNodeWork, avgSecondaryNodeUnique, avgSecondaryLeafWork, "
                "avgSecondaryLeafUnique, avgSecondaryRayTermination, secondaryNodeWorkMax, secondaryNodeWorkMin, "
                "secondaryLeafWorkMax, secondaryLeafWorkMin\n")
        cum_p = cum_s = np.float32(0)
        for j in range(max_size):
            p = list(_step_averages(primary, j, n_tiles_f, wgsq))
            s = list(_step_averages(secondary, j, n_tiles_f, wgsq))
            if j:
                p[4] = np.float32(p[4] + cum_p)
                s[4] = np.float32(s[4] + cum_s)
            cum_p, cum_s = p[4], s[4]
            cells = [str(j)] + [_g(v) for v in p[:5]] + [str(v) for v in p[5:]] + \
                [_g(v) for v in s[:5]] + [str(v) for v in s[5:]]
            f.write(", ".join(cells) + "\n")
    pn, pl = (np.asarray(a, np.int64) for a in primary_visits)
    sn, sl = (np.asarray(a, np.int64) for a in secondary_visits)
    with open(out[3], "w") as f:
        f.write("loadedPrimaryNodes, loadedPrimaryLeafs, loadedPrimaryNodesMax, loadedPrimaryLeafsMax, "
                "loadedPrimaryNodesMin, loadedPrimaryLeafsMin, loadedSecondaryNodes, loadedSecondaryLeafs, "
                "loadedSecondaryNodesMax, loadedSecondaryLeafsMax, loadedSecondaryNodesMin, "
                "loadedSecondaryLeafsMin, loadedWidePrimaryNodes, loadedWidePrimaryLeafs, "
                "loadedWideSecondaryNodes, loadedWideSecondaryLeafs\n")

        def mm(v):
            # max from 0, min from 1000 (reset to 0 when it stays 1000)
            mx, mn = max(0, int(v.max())), min(1000, int(v.min()))
            return mx, (0 if mn == 1000 else mn)

        for i in range(n_tiles):
            sl_ = slice(i * wgsq, (i + 1) * wgsq)
            cells = []
            for nodes, leaves in ((pn[sl_], pl[sl_]), (sn[sl_], sl[sl_])):
                nmx, nmn = mm(nodes)
                lmx, lmn = mm(leaves)
                cells += [_g(np.float32(nodes.sum())), _g(np.float32(leaves.sum())), str(nmx), str(lmx),
                          str(nmn), str(lmn)]
            cells += [_g(np.float32(int(primary[i]["node_unique"].sum()))),
                      _g(np.float32(int(primary[i]["leaf_unique"].sum()))),
                      _g(np.float32(int(secondary[i]["node_unique"].sum()))),
                      _g(np.float32(int(secondary[i]["leaf_unique"].sum())))]
            f.write(", ".join(cells) + "\n")
    return out


def workgroup_analysis(scene, bvh, pos, target, width, height, wg, ambient_samples, ambient_distance, path,
                       name, problem, camera_id=0, wide_alternative=True):
    """The wide render of CameraData::renderImage with doWorkGroupAnalysis
    (cameraData.h:458-560): for every full wg x wg tile, the per-step
    workload of its primary rays (closest hit) and of its ambient rays (any
    hit; a pixel without a primary hit gets a NaN ray that is not started),
    one batch per ambient sample; then the four analysis files.  Returns
    (primary, secondary) per-tile records and the file paths."""
    if ambient_samples < 1:
        raise ValueError("the workgroup analysis needs at least one ambient sample")
    wgsq = wg * wg
    n_tiles = (width // wg) * (height // wg)
    rays = camera_rays(pos, target, width, height, wg)[:n_tiles * wgsq]
    primary, pc = scene.workgroup_steps(rays, wgsq, mode="closest")
    hits = scene.intersect(rays)
    nrm, spos = bvh.surface(rays, hits)
    secondary = [[] for _ in range(n_tiles)]
    sn = np.zeros(len(rays), np.int64)
    sl = np.zeros(len(rays), np.int64)
    for a in range(ambient_samples):
        # a pixel without a hit gets a NaN ray (cameraData.h:495-499)
        amb = np.zeros(len(rays), RAY_DTYPE)
        for f in RAY_DTYPE.names:
            amb[f] = np.float32(np.nan)
        sub, src = spawn_ambient(nrm, spos, hits, a, ambient_distance)
        amb[src] = sub
        recs, c = scene.workgroup_steps(amb, wgsq, mode="any")
        for i in range(n_tiles):
            secondary[i].append(recs[i])
        sn += c["nodes"]
        sl += c["leaves"]
    secondary = [np.concatenate(t) if t else np.zeros(0, primary[0].dtype) for t in secondary]
    files = write_workgroup_analysis(path, name, problem, camera_id, width, height, wg, primary, secondary,
                                     (pc["nodes"], pc["leaves"]), (sn, sl), wide_alternative)
    return primary, secondary, files


def _hist_add(a, b):
    if a is None:
        return {k: (np.array(v, np.int64) if hasattr(v, "__len__") else int(v)) for k, v in b.items()}
    for k, v in b.items():
        a[k] = a[k] + (np.asarray(v, np.int64) if hasattr(v, "__len__") else int(v))
    return a


def analysis_render(scene, bvh, positions, targets, width, height, wg, ambient_samples, ambient_distance, path,
                    name, problem, wide_render=False, workgroup_analysis_on=False, wide_alternative=True):
    """CameraData's analysis renderer (cameraData.h:84-560) on the GPU: per
    camera the primary rays (closest hit) and ambient_samples ambient rays per
    hit through the instrumented traversal (rt_trace_histograms), their
    counters summed over the cameras into <path>/<name><problem>_Info.txt;
    with wide_render and workgroup_analysis_on the workgroup files per camera.
    Returns the _Info.txt path."""
    os.makedirs(path, exist_ok=True)
    hp = hs = None
    for c in range(len(positions)):
        rays = camera_rays(positions[c], targets[c], width, height, wg)
        hits, st = scene.histograms(rays)
        hp = _hist_add(hp, st)
        nrm, pos = bvh.surface(rays, hits)
        for a in range(ambient_samples):
            amb, _ = spawn_ambient(nrm, pos, hits, a, ambient_distance)
            if len(amb):
                hs = _hist_add(hs, scene.histograms(amb, mode="any")[1])
        if wide_render and workgroup_analysis_on:
            workgroup_analysis(scene, bvh, positions[c], targets[c], width, height, wg, ambient_samples,
                               ambient_distance, path, name, problem, camera_id=c, wide_alternative=wide_alternative)
    info = os.path.join(path, f"{name}{problem}_Info.txt")
    write_info(info, name, bvh.branching, bvh.leaf_size, hp, hs, bvh_depth=bvh.max_depth + 1)
    return info


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
        ambient = ambient_distance(tris)
        cam = CameraFast(positions, targets, cfg.get("xRes", 1920), cfg.get("yRes", 1088),
                         cfg.get("workGroupSize", 16))
        os.makedirs(self.out_dir, exist_ok=True)
        problem = f"_b{branching}_l{leaf_size}"
        samples = max(0, cfg.get("ambientSampleCount", 1))
        if cfg.get("renderAnalysisImage"):
            # rayTracer.cpp:695-742; the workgroup analysis needs the wide
            # renderer (cameraData.h:121-133)
            analysis_render(scene, bvh, positions, targets, cam.width, cam.height, cam.wg, samples, ambient,
                            self.out_dir, name, problem, wide_render=cfg.get("wideRender", False),
                            workgroup_analysis_on=cfg.get("doWorkGroupAnalysis", False),
                            wide_alternative=cfg.get("wideAlternative", True))
        if not cfg.get("doPerformanceTest", True):
            scene.close()
            return 0.0
        mode = "save_distance" if cfg.get("saveDistance") else "closest"
        # cameraFast.h:134-147: <path>/WorkGroupSize_<wg>_<Normal|Wide|WideOld>/
        kind = ("Wide" if cfg.get("wideAlternative", True) else "WideOld") if cfg.get("wideRender") else "Normal"
        perf_dir = os.path.join(self.out_dir, f"WorkGroupSize_{cam.wg}_{kind}")
        os.makedirs(perf_dir, exist_ok=True)
        # leaf memory (outer) and node memory (inner) up to 3 gangs above the
        # minimum with the memory tests (rayTracer.cpp:41-52,630-640), at most 16
        g = self.gang
        lms = range(bvh.leaf_memory, (min(16, bvh.leaf_memory + 3 * g) if cfg.get("doLeafMemoryTest") else
                                      bvh.leaf_memory) + 1, g)
        nms = range(bvh.node_memory, (min(16, bvh.node_memory + 3 * g) if cfg.get("doNodeMemoryTest") else
                                      bvh.node_memory) + 1, g)
        base_total = None
        for lm in lms:
            for nm in nms:
                base = lm == bvh.leaf_memory and nm == bvh.node_memory
                b2, s2 = bvh, None
                if not base:
                    b2 = Bvh(tris, branching, leaf_size, gang=g, sort_each_split=cfg.get("sortEachSplit", True),
                             leaf_split_option=cfg.get("leafSplitOption", 1), sah_factor=sah, node_memory=nm,
                             leaf_memory=lm)
                    s2 = Scene(b2)
                total, primary = cam.render_images(s2 or scene, b2, samples, ambient, mode=mode,
                                                   save_image=cfg.get("saveImage", False) and base,
                                                   path=self.out_dir, name=name)
                prefix = f"_mb{b2.node_memory}_ml{b2.leaf_memory}"
                write_perf(os.path.join(perf_dir, f"{name}{problem}{prefix}_Perf.txt"), name, problem, g,
                           b2.node_memory, b2.leaf_memory, branching, leaf_size, total, primary, 0.0,
                           b2.average_depth, b2.triangle_count)
                if s2 is not None:
                    s2.close()
                if base:
                    base_total = total
        scene.close()
        return base_total

    def run(self, scenes=("spheres",)):
        results = {}
        for name in scenes:
            sc = SCENES[name]
            tris = sc["make"]()
            pos, tgt = sc["cam"]
            for leaf, branch in sweep(self.cfg):
                results[(name, branch, leaf)] = self.render(name, tris, branch, leaf, [pos], [tgt])
        return results
