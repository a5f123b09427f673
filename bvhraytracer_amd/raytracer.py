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

The file formats and the config.txt parser have one implementation, the C++
renderer layer (include/rt_renderer.hpp); the functions here call it through
include/rt_report.h.
"""
import ctypes
import os
import time

import numpy as np

from .host import (Bvh, Scene, camera_rays, interior_scene, load_sah_factor, spawn_ambient, spheres_scene)
from ._lib import RAY_DTYPE, STEP_DTYPE, RtTraceHist, check, lib


def read_config(path="config.txt"):
    """RayTracer::readConfig over the reference's defaults: for every
    non-empty line that does not start with '#', every key that occurs
    anywhere in the line (substring match) takes the value after the first
    '='; booleans look for "true" / "false" anywhere in the line (neither:
    false, with a warning).  Returns every setting plus "scenarios",
    "cacheSizes" and "_warnings" (what the reference prints)."""
    n = ctypes.c_size_t()
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        rc = lib().rt_report_read_config(os.fsencode(path), buf, cap, ctypes.byref(n))
        if rc != 0 and n.value + 1 > cap:
            cap = n.value + 1
            continue
        check(rc)
        break
    cfg, warnings = {}, []
    for line in buf.value.decode().splitlines():
        k, v = line.split("=", 1)
        if k == "warning":
            warnings.append(v)
        elif k in ("scenarios", "cacheSizes"):
            cfg[k] = [int(x) for x in v.split(",")] if v else []
        elif v in ("true", "false"):
            cfg[k] = v == "true"
        else:
            cfg[k] = int(v)
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
    row (whose tiles hold height % wg rows).  Values of any dtype and shape
    (n, ...) move as whole pixels."""
    values = np.ascontiguousarray(values)
    if wg <= 0 or width % wg:
        raise ValueError("width must be a multiple of the workgroup size")
    if len(values) != width * height:
        raise ValueError(f"expected {width * height} pixels, got {len(values)}")
    raw = values.view(np.uint8).reshape(width * height, -1) if values.size else values.reshape(0, 1)
    out = np.empty_like(values)
    check(lib().rt_report_tile_order_to_rows(raw.ctypes.data, width, height, wg, raw.shape[1],
                                             out.view(np.uint8).ctypes.data))
    return out.reshape((height, width) + values.shape[1:])


def write_png(path, rgba):
    """An 8-bit RGBA PNG of a (height, width, 4) uint8 image (lodepng::encode
    of the reference's encodeTwoSteps, util.h:52-61, takes the same RGBA
    buffer; the pixels decode identically, the compressed bytes differ)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("expected a (height, width, 4) image")
    check(lib().rt_report_write_png(os.fsencode(path), rgba.ctypes.data, rgba.shape[1], rgba.shape[0]))


def ambient_distance(tris):
    """ambientDistance of RayTracer::renderImage (rayTracer.cpp:607): cbrt of
    the root box volume (Aabb::getVolume, aabb.h:136-140) / 10, in float32,
    the root box being the triangles' bounds."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    out = ctypes.c_float()
    check(lib().rt_report_ambient_distance(t.ctypes.data, len(t), ctypes.byref(out)))
    return float(out.value)


def ambient_image(occluded, hit, samples):
    """Gray value per pixel (cameraFast.h:404-439): a hit pixel is
    (uint8)((1 - occluded / samples) * 255) in float32, a miss is 0.  With
    no samples the reference converts NaN, which gives 0 on x86."""
    hit = np.ascontiguousarray(hit, np.uint8)
    occ = np.ascontiguousarray(occluded, np.uint32)
    img = np.zeros(len(hit), np.uint8)
    check(lib().rt_report_ambient_gray(hit.ctypes.data, occ.ctypes.data, len(hit), int(samples), img.ctypes.data))
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
    """The _Perf.txt fields of cameraFast.h:147-161 (`problem` names the
    file, the text does not hold it)."""
    check(lib().rt_report_write_perf(os.fsencode(path), name.encode(), branching, leaf_size, gang, node_memory,
                                     leaf_memory, float(total), float(ray_sum), float(tri_sum), float(avg_depth),
                                     int(tri_count)))


def _hist_struct(st):
    """rt_trace_hist from the dict of Scene.histograms (summed or not)."""
    h = RtTraceHist()
    for f, _ in RtTraceHist._fields_:
        v = st[f]
        if hasattr(v, "__len__"):
            arr = getattr(h, f)
            vals = [int(x) for x in v]
            if len(vals) > len(arr):
                raise ValueError(f"{f}: {len(vals)} bins, the histogram holds {len(arr)}")
            for i, x in enumerate(vals):
                arr[i] = x
        else:
            setattr(h, f, int(v))
    return h


def write_info(path, name, branching, leaf_size, primary, secondary, bvh_depth=None):
    """The _Info.txt of the reference's instrumented renderer
    (cameraData.h:241-329), from rt_trace_histograms counters of the primary
    rays and of the secondary (ambient) rays: per-ray averages, success
    ratios, the waste factor, child fullness 0..N and leaf fullness 0..L of
    the primary rays, and visits per depth 0..bvh_depth-1 (bvhDepth, the
    tree depth + 1; by default up to the deepest primary visit), in float32
    with the reference's line keys and number formats, so its
    Analysis/DataManager.py reads it.  Every ray, test and visit count must
    be nonzero (the reference divides by them)."""
    if bvh_depth is None:
        nz = np.nonzero(np.asarray(primary["node_depth"]) + np.asarray(primary["leaf_depth"]))[0]
        bvh_depth = int(nz[-1]) + 1 if len(nz) else 1
    hp, hs = _hist_struct(primary), _hist_struct(secondary)
    check(lib().rt_report_write_info(os.fsencode(path), name.encode(), branching, leaf_size, int(bvh_depth),
                                     ctypes.byref(hp), ctypes.byref(hs)))


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
    n_tiles = (width // wg) * (height // wg) if wg else 0
    if len(primary) != n_tiles or len(secondary) != n_tiles:
        raise ValueError(f"expected {n_tiles} tiles")

    def flat(tiles):
        recs = [np.ascontiguousarray(t, STEP_DTYPE) for t in tiles]
        lens = np.array([len(t) for t in recs], np.uint32)
        cat = np.concatenate(recs) if recs else np.zeros(0, STEP_DTYPE)
        return np.ascontiguousarray(cat), lens

    p, pl = flat(primary)
    s, sl = flat(secondary)
    nrays = n_tiles * wg * wg
    vis = []
    for a in (*primary_visits, *secondary_visits):
        v = np.ascontiguousarray(a, np.int64)
        if len(v) < nrays or (len(v) and (v.min() < 0 or v.max() > 0xFFFFFFFF)):
            raise ValueError(f"expected {nrays} visit counts per ray in uint32 range")
        vis.append(np.ascontiguousarray(v[:nrays], np.uint32))
    cap = 4 * (len(os.fsencode(path)) + len(name.encode()) + len(problem.encode()) + 256)
    buf = ctypes.create_string_buffer(cap)
    check(lib().rt_report_write_workgroup_analysis(
        os.fsencode(path), name.encode(), problem.encode(), int(camera_id), width, height, wg, p.ctypes.data,
        pl.ctypes.data, s.ctypes.data, sl.ctypes.data, vis[0].ctypes.data, vis[1].ctypes.data, vis[2].ctypes.data,
        vis[3].ctypes.data, int(bool(wide_alternative)), buf, cap))
    return buf.value.decode().splitlines()


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
