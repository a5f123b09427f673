"""Per-scene autotuner: N, L, node format, workgroup size, LDS stack window and
wave schedule.

The reference picks N and L by sweeping them offline and keeps the best SAH
node factor per (N, L) in a table (sahFactorConfigSse.txt, loaded by
RayTracer::loadSahFactor, rayTracer.cpp:880-925).  Here the sweep is done per
scene on the GPU: every candidate (N, L) is built with the reference builder
and uploaded, and every launch shape is timed on a sample of the scene's own
rays; the fastest configuration wins.  Results are cached by a content hash
of the triangles and the ray sample, so a scene is measured once.
"""
import hashlib
import json
import os
import time

import numpy as np

from .host import Bvh, Scene

DEFAULT_NL = ((4, 4), (8, 8), (8, 4), (16, 8), (4, 8), (16, 16))
DEFAULT_BLOCKS = (64, 128, 256)
DEFAULT_WINDOWS = (8, 16, 32)
DEFAULT_SCHEDULES = (0, 8, 16, 24)  # rt_set_schedule leaf ratios (x16)
# node formats: the quantized records (conservative 8-bit child planes, the
# default) and the exact-box records (RT_EXACT_BOXES); which is faster depends
# on the rays (the quantized ones move fewer bytes, the exact ones issue fewer
# instructions, which coherent rays favour)
DEFAULT_FORMATS = (False, True)  # exact_boxes


def scene_key(tris, rays):
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(tris).view(np.uint8))
    h.update(np.ascontiguousarray(rays).view(np.uint8))
    return h.hexdigest()


class Autotuner:
    def __init__(self, cache_path=None, nl=DEFAULT_NL, blocks=DEFAULT_BLOCKS, windows=DEFAULT_WINDOWS,
                 schedules=DEFAULT_SCHEDULES, sample=1 << 16, reps=3, fill=1 << 20, formats=DEFAULT_FORMATS):
        """sample: distinct rays of the scene timed; fill: rays per timed
        launch (the sample repeated), enough to fill the GPU so the timing
        measures throughput rather than the latency of the longest ray;
        formats: the exact_boxes values tried."""
        self.cache_path = cache_path
        self.nl, self.blocks, self.windows = tuple(nl), tuple(blocks), tuple(windows)
        self.schedules = tuple(schedules)
        self.formats = tuple(formats)
        self.sample, self.reps, self.fill = int(sample), int(reps), int(fill)
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def save(self):
        if self.cache_path:
            os.makedirs(os.path.dirname(os.path.abspath(self.cache_path)), exist_ok=True)
            with open(self.cache_path, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)

    def _time(self, scene, d_rays, d_out, mode, exact_boxes=False):
        best = float("inf")
        for _ in range(self.reps):
            scene.intersect_device(d_rays, d_out, mode=mode, exact_boxes=exact_boxes)
            best = min(best, scene.stats()["last_kernel_ms"])
        return best

    # node format, block size, stack window, schedule of a new scene (the
    # quantized records' default shape)
    DEFAULT_LAUNCH = (False, 128, 8, 16)

    def tune_launch(self, scene, d_rays, d_out, mode="closest"):
        """Fastest (node format, block size, stack window, wave schedule) for
        this scene and ray set; the scene keeps it (Scene.exact_boxes,
        set_launch, set_schedule).  default_ms is the time of the default
        launch (None when it is not among the candidates)."""
        res = []
        formats = self.formats if scene.stats()["quantized_record_bytes"] else (True,)
        for ex in formats:
            for w in self.windows:
                for b in self.blocks:
                    for sc in self.schedules:
                        scene.set_launch(b, w)
                        scene.set_schedule(sc)
                        res.append((self._time(scene, d_rays, d_out, mode, ex), ex, b, w, sc))
        ms, ex, b, w, sc = min(res)
        scene.set_launch(b, w)
        scene.set_schedule(sc)
        scene.exact_boxes = ex
        default = [r[0] for r in res if r[1:] == self.DEFAULT_LAUNCH]
        return {"exact_boxes": ex, "block": b, "window": w, "schedule": sc, "ms": ms,
                "default_ms": default[0] if default else None}

    def tune_scene(self, tris, rays, mode="closest", key=None):
        """Returns {"N", "L", "exact_boxes", "block", "window", "schedule",
        "mrays", "tried"} for the scene.
        rays: host RAY_DTYPE array; a sample of `sample` rays is used."""
        import torch

        sample = rays[: self.sample]
        key = key or scene_key(tris, sample) + f":{mode}:{self.fill}"
        if key in self.cache:
            return dict(self.cache[key], cached=True)
        timed = np.resize(np.ascontiguousarray(sample), max(len(sample), self.fill))
        d_rays = torch.from_numpy(np.ascontiguousarray(timed).view(np.float32).reshape(-1, 10)).cuda()
        if mode == "any":
            d_out = torch.empty(len(timed), dtype=torch.uint8, device="cuda")
        else:
            d_out = torch.empty((len(timed), 6), dtype=torch.int32, device="cuda")
        t0 = time.time()
        best, tried = None, []
        for N, L in self.nl:
            bvh = Bvh(tris, N, L)
            scene = Scene(bvh)
            r = self.tune_launch(scene, d_rays, d_out, mode)
            mrays = len(timed) / r["ms"] / 1e3
            cfg = {"N": N, "L": L, "exact_boxes": r["exact_boxes"], "block": r["block"], "window": r["window"],
                   "schedule": r["schedule"],
                   "mrays": round(mrays, 2),
                   "default_mrays": round(len(timed) / r["default_ms"] / 1e3, 2) if r["default_ms"] else None}
            tried.append(cfg)
            if best is None or mrays > best["mrays"]:
                best = dict(cfg)
            scene.close()
            bvh.close()
        best["tried"] = tried
        best["seconds"] = round(time.time() - t0, 2)
        self.cache[key] = best
        return dict(best, cached=False)
