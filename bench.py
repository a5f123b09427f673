"""Benchmark: Mrays/s of the closest-hit traversal on one MI355X.

Workload (the metric's configuration): the 1M-triangle clustered interior
scene, SAH BVH at N8L8 (gang 4), 1920x1080 primary rays from (0,5,-14) to
(0,3,0); their hits seed 8,000,000 cosine-distributed diffuse bounce rays
(seed 3).  One step = one rt_trace of those 8M incoherent closest-hit rays
with rays and results resident in HBM.

Prints one JSON line (see the contract in the repository docs); the CPU
baseline is the oracle (C restatement of the reference's ISPC kernels and
FastNodeManager traversal) on a bounded sample of the same rays, run on the
host cores with OpenMP dynamic(4) over 256-ray tiles like CameraFast.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0

# HBM bytes per launch of the timed trace kernel, from two rocprofv3 --pmc
# runs of this benchmark (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
# each scaled by what the same counter reads for a plain 1 GiB device copy
# (bench.py --calibrate-copy): bytes = FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0,
# in KiB.  Keyed by (branching, leaf size, rays); the commands and the
# per-dispatch values are in profiles/k_trace_quantized_N8L8_diffuse8M_counters.md
# and profiles/k_trace_batch_quantized_64scenes.md (the calibration in
# profiles/k_trace_exact_N8L8_diffuse8M_hbm_traffic.md).
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_TRAFFIC_KIB = {(8, 8, 8_000_000): {"fetch": 6776918, "write": 742906},
                        ("batch", 8, 8, 9_872_996): {"fetch": 2318751, "write": 501051}}

# What the timed kernel spends, from rocprofv3 counter runs of this benchmark
# (profiles/k_trace_quantized_N8L8_diffuse8M_counters.md): 573 VALU and 12.5
# load instructions per ray at 41% active lanes and 6 waves per SIMD; the
# texture address unit 70% busy and the data unit 95% (55% of it stalled on
# the L1), 3.44e8 L1-to-L2 requests (22 GB) per step, 68% L2 hits, and HBM
# carrying about a quarter of its peak.
MEASURED_COUNTERS = {(8, 8, 8_000_000): {"td_busy": 0.955, "td_stalled": 0.551, "ta_busy": 0.696,
                                         "valu_busy_metric": 1.155, "active_lanes": 0.414, "l2_hit": 0.678,
                                         "l2_read_bytes": 22_007_040_000,
                                         "valu_insts_per_ray": 573.2, "loads_per_ray": 12.46,
                                         "source": "profiles/k_trace_quantized_N8L8_diffuse8M_counters.md"}}


def limiter_text(c):
    """What binds the kernel, from the counter runs."""
    if not c or "td_busy" not in c:
        return None
    return (f"the load pipeline: one line per lane for every record load, whose rate the gather probe "
            f"bounds (gather.frac; TD {c['td_busy']:.0%} busy, {c['td_stalled']:.0%} stalled on the L1; "
            f"TA {c['ta_busy']:.0%}), beside VALU issue ({c['valu_insts_per_ray']:.0f} instructions per ray at "
            f"{c['active_lanes']:.0%} active lanes; the VALUBusy metric reads {c['valu_busy_metric']:.0%}); "
            f"HBM carries a fraction of its peak (hbm_frac)")


def gather_roofline(c, rays, k_ms):
    """The kernel's record loads against the load pipeline's ceiling: its
    wave-level loads per second (the counter runs' loads per ray x rays over
    this run's kernel time) over rt_measure_gather's rate when every active
    lane loads from its own 128-byte line (L1-resident, 6 waves per SIMD), at
    the kernel's VALU active-lane fraction and at a full wave."""
    from bvhraytracer_amd.host import measure_gather
    if not c or "loads_per_ray" not in c:
        return None
    lanes = max(1, min(64, int(round(c["active_lanes"] * 64))))
    peak = measure_gather(lanes)
    full = measure_gather(64)
    achieved = c["loads_per_ray"] * rays / (k_ms / 1000.0)
    return {"unit": "wave loads/s", "achieved": float(f"{achieved:.4g}"), "peak": float(f"{peak:.4g}"),
            "active_lanes": lanes, "frac": round(achieved / peak, 4),
            "peak_64_lanes": float(f"{full:.4g}"), "frac_64_lanes": round(achieved / full, 4),
            "probe": "rt_measure_gather: 16-byte loads, one 128-byte line per lane, 16 KB buffer"}


def measured_traffic(key):
    mt = MEASURED_TRAFFIC_KIB.get(key)
    return None if mt is None else int((mt["fetch"] * FETCH_SCALE + mt["write"] * WRITE_SCALE) * 1024)
METRIC = "Mrays/s per MI355X (closest-hit, N8L8) batched vs plain; p99 request latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def header_bytes(M):
    # inner-node header bytes the kernel loads per visit (rt_layout.h): one
    # 16-byte block for M <= 8, three for M = 12, 16
    return 16 if M <= 8 else 48


def algorithmic_bytes(counts, M):
    """Bytes the traversal must move per the device layout: ray in (40 B),
    hit out (24 B), 24*k box bytes + header per inner visit, 36*k per leaf
    (the count of slots to test rides in the leaf's first load)."""
    c = counts
    return (64 * len(c) + 24 * c["node_slots"].astype(np.float64).sum()
            + header_bytes(M) * c["nodes"].astype(np.float64).sum()
            + 36 * c["leaf_slots"].astype(np.float64).sum())


def quantized_plane_bytes(M):
    # bytes of quantized child planes a quantized inner record visit loads
    # (rt_layout.h): 2, 3 or 6 sixteen-byte blocks
    return 32 if M == 4 else (48 if M == 8 else 96)


def layout_bytes(counts, M):
    """Bytes the quantized node records move: ray in and hit out (64 B), per
    inner visit the own box and frame (32 B), the header block (16 B, 48 B
    for M > 8) and the quantized planes; per leaf visit the 32-byte head and
    40 B per slot (a pair of triangles is one 80-byte chunk).  Tentative pops
    whose exact box test fails (2-3% of the pops on these rays) are not
    counted."""
    c = counts
    return (64 * len(c) + (32 + header_bytes(M) + quantized_plane_bytes(M)) * c["nodes"].astype(np.float64).sum()
            + 32 * c["leaves"].astype(np.float64).sum() + 40 * c["leaf_slots"].astype(np.float64).sum())


def useful_bytes(n, hist, M):
    """Bytes of real work only: ray in and hit out (64 B per ray), the real
    child boxes tested (24 B each) plus the header per inner visit, and the
    real triangles tested (36 B each); the padded slots that the loop
    counts k include are left out.  hist: rt_trace_histograms totals."""
    return (64.0 * n + 24.0 * hist["aabb_tests"] + header_bytes(M) * hist["node_visits"]
            + 36.0 * hist["tri_tests"])


def hbm_frac(traffic, kernel_ms):
    return None if traffic is None else round(traffic / (kernel_ms / 1000.0) / 1e9 / HBM_PEAK_GBPS, 4)


def median_wall(fn, torch, reps, warmup=1):
    """Median wall time (s) of fn() followed by a device synchronize."""
    for _ in range(max(1, warmup)):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def bench_batched(rt, torch, args, log):
    """64 scenes (seeds 100..163), triangles clamp(LogNormal(ln 2e5, 1), 1e4, 2e6),
    rays per scene clamp(LogNormal(ln 6.4e4, 1.2), 4096, 2^20), half primary
    (random camera in the room), half diffuse bounces from those hits.
    batched = one rt_batch_trace (captured graph, re-armed); plain_sync = one
    rt_trace per scene that waits for its kernel (the call's default);
    plain_async = the same 64 rt_trace calls queued back to back on one
    stream without waiting (RT_NO_SYNC), one synchronize at the end."""
    from bvhraytracer_amd import workloads as wl

    N, L = args.branching, args.leaf
    ntris, nrays = wl.batch_sizes()
    t0 = time.time()
    scenes = []
    for i in range(wl.BATCH_SCENES):
        tris, bvh, sc, rays = wl.batch_scene(i, N, L, ntris[i], nrays[i])
        d_r = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
        d_h = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
        scenes.append((tris, bvh, sc, rays, d_r, d_h))
    total = sum(len(x[3]) for x in scenes)
    log(f"batched workload: 64 scenes, {int(ntris.sum())} tris, {total} rays, setup {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()

    def plain_sync():
        for x in scenes:
            x[2].intersect_device(x[4], x[5], stream=stream)

    def plain_async():
        for x in scenes:
            x[2].intersect_device(x[4], x[5], stream=stream, sync=False)

    batch = rt.Batch(64)
    outs = [torch.empty_like(x[5]) for x in scenes]
    items = [(x[2], x[4], o, "closest") for x, o in zip(scenes, outs)]

    def batched():
        batch.trace(items, graph=True, stream=stream, sync=False)

    reps = max(3, args.steps // 2)
    res = {name: total / median_wall(fn, torch, reps, args.warmup) / 1e6
           for name, fn in (("plain_sync", plain_sync), ("plain_async", plain_async), ("batched", batched))}
    same = all(torch.equal(o, x[5]) for o, x in zip(outs, scenes))
    # kernel time of the batched launch (HIP events around the graph launch)
    kms = []
    for _ in range(reps):
        batch.trace(items, graph=True, stream=stream)
        kms.append(batch.stats()["last_kernel_ms"])
    k_ms = float(np.median(kms))
    # algorithmic bytes of the same launch from a counts pass
    cnts = [torch.empty((len(x[3]), 4), dtype=torch.int32, device="cuda") for x in scenes]
    batch.trace([(x[2], x[4], o, "closest", c) for x, o, c in zip(scenes, outs, cnts)], graph=True, stream=stream)
    host_cnts = [c.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1) for c in cnts]
    alg = sum(algorithmic_bytes(c, x[1].node_memory) for x, c in zip(scenes, host_cnts))
    lay = sum(layout_bytes(c, x[1].node_memory) for x, c in zip(scenes, host_cnts))
    achieved = alg / (k_ms / 1000.0) / 1e9
    traffic = measured_traffic(("batch", N, L, total))
    log(f"batched {res['batched']:.1f} Mrays/s (kernel {k_ms:.3f} ms, {achieved:.0f} GB/s algorithmic) vs "
        f"plain {res['plain_sync']:.1f} waiting / {res['plain_async']:.1f} queued, identical={same}")
    batch.close()
    out = {"scenes": 64, "triangles": int(ntris.sum()), "rays": int(total),
           "plain_sync_mrays": round(res["plain_sync"], 2), "plain_async_mrays": round(res["plain_async"], 2),
           "batched_mrays": round(res["batched"], 2),
           "speedup_vs_sync": round(res["batched"] / res["plain_sync"], 3),
           "speedup_vs_async": round(res["batched"] / res["plain_async"], 3),
           "identical_to_plain": bool(same), "n_l": f"N{N}L{L}", "graph": True,
           "kernel_ms": round(k_ms, 4), "kernel_mrays": round(total / k_ms / 1e3, 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                        "hbm_frac": hbm_frac(traffic, k_ms), "algorithmic_bytes": int(alg),
                        "layout_bytes": int(lay), "layout_frac": round(lay / (k_ms / 1000.0) / 1e9 / HBM_PEAK_GBPS, 4)}}
    if not args.no_batch_autotune:
        out["autotuned"] = batched_autotuned(rt, torch, scenes, outs, stream, total, args, log)
    if not args.no_cpu:
        # the CPU baseline of this workload: the SIMD oracle per scene,
        # summed (one warm-up pass, then one timed pass over all 64 scenes)
        import oracle

        orcs = [(oracle.Oracle.from_bvh(x[1], simd=True), x[3]) for x in scenes]
        for o, rays in orcs:
            o.trace(rays, counts=False)
        t = time.perf_counter()
        for o, rays in orcs:
            used = o.trace(rays, counts=False)["threads"]
        cpu_s = time.perf_counter() - t
        out["cpu_port_mrays"] = round(total / cpu_s / 1e6, 3)
        out["cpu_threads"] = int(used)
        log(f"batched workload on the CPU oracle: {out['cpu_port_mrays']} Mrays/s on {used} threads")
    for x in scenes:
        x[2].close()
    return out


BATCH_SHAPES = ((128, 8), (128, 16), (128, 32), (256, 8), (256, 16), (256, 32))


def batched_autotuned(rt, torch, scenes, ref_outs, stream, total, args, log):
    """The same 64 scenes with N and L chosen per scene by the autotuner
    (candidates N4L4, N4L8, N8L4, N8L8, each with both node formats, block
    size 128/256 and LDS stack window 8/16, timed on 8192 of the scene's own
    rays repeated to
    2^20 per launch, so the GPU is full), traced in one batched graph launch
    that mixes the chosen node and leaf memories.  Block size and stack
    window are properties of a launch (its blocks and its LDS allocation),
    and so is the node format (one kernel per launch).  Two ways to honour
    them: one format and shape for the whole launch (the fastest of the
    quantized and exact-box records at BATCH_SHAPES), or every scene with its
    own tuned format, shape and schedule (RT_SCENE_SHAPES: one persistent
    launch per distinct shape, in one graph).  The tuner's last step compares
    these with N8L8 for every scene, at each format and shape, and keeps the
    fastest (per-scene timings alone do not see how the scenes share the GPU
    in one launch).  The per-scene shape choices are reported with what they
    gain over the default shape on the scene alone.  Tuning and the rebuilds
    are ingest cost, outside the timing."""
    from bvhraytracer_amd.autotune import Autotuner

    t0 = time.time()
    tuner = Autotuner(nl=((4, 4), (4, 8), (8, 4), (8, 8)), blocks=(128, 256), windows=(8, 16),
                      schedules=(16,), sample=8192, reps=2, fill=1 << 20)
    tuned, choice, shapes, gains, bests = [], {}, {}, [], []
    for tris, bvh0, sc0, rays, d_r, d_h in scenes:
        best = tuner.tune_scene(tris, rays)
        bests.append(best)
        key = f"N{best['N']}L{best['L']}"
        choice[key] = choice.get(key, 0) + 1
        shape = f"{'x' if best['exact_boxes'] else 'q'}{best['block']}x{best['window']}"
        shapes[shape] = shapes.get(shape, 0) + 1
        if best.get("default_mrays"):
            gains.append(best["mrays"] / best["default_mrays"])
        if (best["N"], best["L"]) == (bvh0.branching, bvh0.leaf_size):
            tuned.append((None, sc0, d_r, torch.empty_like(d_h)))
        else:
            bvh = rt.Bvh(tris, best["N"], best["L"])
            tuned.append((bvh, rt.Scene(bvh), d_r, torch.empty_like(d_h)))
    tune_s = time.time() - t0
    batch = rt.Batch(64)
    tuned_items = [(sc, d_r, o, "closest") for (bvh, sc, d_r, o) in tuned]
    uniform_items = [(sc0, d_r, torch.empty_like(d_h), "closest") for (tris, bvh0, sc0, rays, d_r, d_h) in scenes]
    reps = max(3, args.steps // 2)

    def by_shape_of(items):
        """Mrays/s by node format and launch shape ("q" quantized records,
        "x" exact-box records)."""
        out = {}
        for ex in (False, True):
            for bd, win in BATCH_SHAPES:
                batch.set_launch(bd, win)
                sec = median_wall(lambda: batch.trace(items, graph=True, stream=stream, sync=False, exact_boxes=ex),
                                  torch, reps, args.warmup)
                out[f"{'x' if ex else 'q'}{bd}x{win}"] = round(total / sec / 1e6, 2)
        return out

    by_shape = by_shape_of(tuned_items)
    uniform_by_shape = by_shape_of(uniform_items)
    # every scene with its own tuned format, shape and schedule
    # (RT_SCENE_SHAPES): one persistent launch per distinct shape, one graph
    for (bvh, sc, d_r, o), best in zip(tuned, bests):
        sc.set_launch(best["block"], best["window"])
        sc.set_schedule(best["schedule"])
        sc.exact_boxes = best["exact_boxes"]
    segments = len({(b["exact_boxes"], b["block"], b["window"], b["schedule"]) for b in bests})
    sec = median_wall(lambda: batch.trace(tuned_items, graph=True, stream=stream, sync=False, scene_shapes=True),
                      torch, reps, args.warmup)
    scene_shapes_mrays = total / sec / 1e6
    shaped_outs = [o.clone() for (b, s_, d, o) in tuned]
    per_scene_best = max(by_shape.values())
    candidates = {"per-scene N/L": per_scene_best, "N8L8 for every scene": max(uniform_by_shape.values()),
                  "per-scene N/L and shapes": scene_shapes_mrays}
    kept = max(candidates, key=candidates.get)
    if kept == "per-scene N/L and shapes":
        best_shape = f"per scene ({segments} launches)"
        run = lambda: batch.trace(tuned_items, graph=True, stream=stream, sync=False, scene_shapes=True)  # noqa: E731
    else:
        items = tuned_items if kept == "per-scene N/L" else uniform_items
        shapes_of = by_shape if kept == "per-scene N/L" else uniform_by_shape
        best_shape = max(shapes_of, key=shapes_of.get)
        exact = best_shape[0] == "x"
        bd, win = (int(v) for v in best_shape[1:].split("x"))
        batch.set_launch(bd, win)
        run = lambda: batch.trace(items, graph=True, stream=stream, sync=False, exact_boxes=exact)  # noqa: E731
    # the kept configuration timed again, with three times the repetitions
    sec = median_wall(run, torch, 3 * reps, args.warmup)
    mrays = total / sec / 1e6
    # the per-scene-shape launches give the same results as one launch shape
    batch.set_launch(128, 8)
    batch.trace(tuned_items, graph=True, stream=stream)
    shapes_identical = all(torch.equal(a, o) for a, (b, s_, d, o) in zip(shaped_outs, tuned))
    # the hit triangles are the same whatever the tree, up to exact ties
    same_tri = sum(int((o[:, 0] == r[:, 0]).sum().item()) for (b, s, d, o), r in zip(tuned, ref_outs)) / total
    log(f"batched autotuned {mrays:.1f} Mrays/s ({kept}, launch shape {best_shape}; per-scene N/L by shape "
        f"{by_shape}, N8L8 by shape {uniform_by_shape}, per-scene N/L and shapes {scene_shapes_mrays:.1f} in "
        f"{segments} launches, identical {shapes_identical}), choices {choice}, per-scene shapes {shapes}, "
        f"tuning {tune_s:.1f}s")
    batch.close()
    for bvh, sc, d_r, o in tuned:
        if bvh is not None:
            sc.close()
    return {"batched_mrays": round(mrays, 2), "kept": kept, "launch_shape": best_shape,
            "per_scene_nl_mrays_by_shape": by_shape, "uniform_n8l8_mrays_by_shape": uniform_by_shape,
            "per_scene_shapes_mrays": round(scene_shapes_mrays, 2), "per_scene_shape_launches": segments,
            "per_scene_shapes_identical": shapes_identical,
            "choices": choice, "per_scene_shapes": shapes,
            "per_scene_shape_gain_median": round(float(np.median(gains)), 4) if gains else None,
            "per_scene_shape_gain_max": round(float(np.max(gains)), 4) if gains else None,
            "tune_seconds": round(tune_s, 1), "same_hit_triangle_fraction": round(same_tri, 6)}


def bench_frame(rt, torch, scene, bvh, n_bounce, log, reps=3):
    """One bounce frame: 1920x1080 primaries -> closest hit -> hit surfaces ->
    n_bounce diffuse rays -> closest hit (camera rays on the exact-box
    records, the bounces on the quantized ones).  device: every stage on the GPU on
    one stream (rt_camera_rays_device, rt_hit_surface_device,
    rt_spawn_diffuse_device, device-resident rt_trace); host-assisted: the
    host makes the rays and surfaces and each trace copies rays in and hits
    out.  Median of reps frames after one warm-up, wall clock."""
    W, H, pos, tgt = 1920, 1080, (0.0, 5.0, -14.0), (0.0, 3.0, 0.0)
    stream = torch.cuda.Stream()
    n = W * H
    d_h = torch.empty((n, 6), dtype=torch.int32, device="cuda")
    d_bh = torch.empty((n_bounce, 6), dtype=torch.int32, device="cuda")

    def device_frame():
        with torch.cuda.stream(stream):
            d_r = rt.camera_rays_device(pos, tgt, W, H, 16, stream=stream)
            scene.intersect_device(d_r, d_h, stream=stream, sync=False, exact_boxes=True)
            d_n, d_p = scene.surface_device(d_r, d_h, stream=stream)
            d_b = rt.spawn_diffuse_device(d_r, d_n, d_p, d_h, n_bounce, seed=3, stream=stream)
            scene.intersect_device(d_b, d_bh, stream=stream, sync=False)
        stream.synchronize()

    def host_frame():
        prim = rt.camera_rays(pos, tgt, W, H, 16)
        hits = scene.intersect(prim, exact_boxes=True)
        nrm, p = bvh.surface(prim, hits)
        bounce = rt.spawn_diffuse(prim, nrm, p, hits, n_bounce, seed=3)
        return scene.intersect(bounce)

    out = {}
    for name, fn in (("device", device_frame), ("host_assisted", host_frame)):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name + "_ms"] = round(float(np.median(ts)) * 1000.0, 3)
    out["rays"] = n + n_bounce
    out["device_mrays"] = round((n + n_bounce) / (out["device_ms"] / 1000.0) / 1e6, 2)
    log(f"bounce frame: device {out['device_ms']} ms, host-assisted {out['host_assisted_ms']} ms")
    return out


def bench_copy(torch, log, nbytes=1 << 30, reps=5):
    """STREAM-like device copy of nbytes (read nbytes + write nbytes), the
    HBM rate a plain streaming kernel reaches on this GPU: median of reps
    after a warm-up, HIP events."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    gbps = 2 * nbytes / (float(np.median(ms)) / 1000.0) / 1e9
    log(f"device copy: {gbps:.0f} GB/s (read + write of {nbytes} bytes)")
    del a, b
    return {"bytes_read_plus_written": 2 * nbytes, "gbps": round(gbps, 1)}


def bench_render_frame(rt, scene, bvh, tris, log, reps=3):
    """One frame of the reference's performance renderer (CameraFast::
    renderImage: 1920x1088 primaries, closest hit, one ambient-occlusion ray
    per hit): device = rt_render_ambient (every stage on the GPU, wall clock
    of the call), host-assisted = the host-driven render_image (host rays and
    surfaces, host-pointer traces).  Median of reps frames after a warm-up;
    the two gray images must be identical."""
    from bvhraytracer_amd.raytracer import CameraFast, ambient_distance

    cam = CameraFast([(0.0, 5.0, -14.0)], [(0.0, 3.0, 0.0)], 1920, 1088, 16)
    dist = ambient_distance(tris)
    out, imgs = {}, {}
    fns = (("device", lambda: scene.render_ambient(cam.positions[0], cam.targets[0], 1920, 1088, 16, 1, dist)[0]),
           ("host_assisted", lambda: cam.render_image(scene, bvh, 0, 1, dist)[3]))
    for name, fn in fns:
        imgs[name] = fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name + "_ms"] = round(float(np.median(ts)) * 1000.0, 3)
    out["identical"] = bool(np.array_equal(imgs["device"], imgs["host_assisted"]))
    log(f"render frame: device {out['device_ms']} ms, host-assisted {out['host_assisted_ms']} ms, "
        f"identical {out['identical']}")
    return out


def serve_run(rt, srv, sids, world, rate, seconds, seed, device=None):
    """Poisson arrivals at `rate` req/s for `seconds` (workloads.serving_requests)
    submitted at their arrival times from this thread while a second thread
    waits for them in order.  Returns the server's statistics of the window
    plus the arrival rate the submitting thread actually reached.  device:
    per scene the (closest-hit rays, shadow rays) as CUDA tensors; the
    requests then stay in device memory (RT_DEVICE_PTRS), results too."""
    import queue
    import threading

    from bvhraytracer_amd import workloads as wl

    gaps, sizes, scene_of, shadow, offs = wl.serving_requests(rate, seconds, seed)
    n_req = len(gaps)
    # result buffers allocated and touched before the window, as a client
    # that reuses its buffers has them (first-touch page faults of fresh
    # allocations would otherwise be timed as part of the server's copies)
    if device is None:
        outs = [np.zeros(int(sizes[k]), np.uint8) if shadow[k] else np.zeros(int(sizes[k]), rt.HIT_DTYPE)
                for k in range(n_req)]
    else:
        import torch
        outs = [torch.zeros(int(sizes[k]), dtype=torch.uint8, device="cuda") if shadow[k]
                else torch.zeros((int(sizes[k]), 6), dtype=torch.int32, device="cuda") for k in range(n_req)]
        torch.cuda.synchronize()
    srv.stats(reset=True)
    done_q = queue.SimpleQueue()

    def waiter():
        for _ in range(n_req):
            rid = done_q.get()
            try:
                srv.wait(rid, 120)  # the server frees its hold on the output buffer
            except rt.RtError:
                pass  # counted as failed by the server

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    at = start
    for i in range(n_req):
        at += gaps[i]
        dt = at - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        w = world[scene_of[i]] if device is None else device[scene_of[i]]
        src = (w[4] if shadow[i] else w[3]) if device is None else (w[1] if shadow[i] else w[0])
        rid, _ = srv.submit(sids[scene_of[i]], src[offs[i]:offs[i] + sizes[i]], "any" if shadow[i] else "closest",
                            out=outs[i])
        done_q.put(rid)
    submit_s = time.perf_counter() - start
    th.join()
    st = srv.stats()
    return {"arrival_rate": rate, "requests": n_req, "submitted_rate": round(n_req / submit_s, 1),
            "offered_mrays_per_s": round(float(sizes.sum()) / max(submit_s, 1e-9) / 1e6, 2),
            "mrays_per_s": round(st["mrays_per_s"], 2), "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
            "requests_per_launch": round(st["mean_requests_per_launch"], 2), "completed": int(st["completed"]),
            "failed": int(st["failed"])}


SWEEP_RATES = (2000, 10000, 20000, 40000)
SWEEP_REQUESTS = 3000
DEVICE_SWEEP_RATES = (40000, 80000)


def bench_serving(rt, args, log):
    """16 resident scenes (100k..2M tris log-spaced, seeds 200..215) at N/L.
    First the configuration itself: Poisson arrivals at 200 req/s for
    args.serve_seconds (seed 5), rays per request clamp(LogNormal(ln 2.3e4,
    1.3), 1024, 2^19), half shadow (any-hit), half closest-hit; latency from
    submit to results in host memory.  Then a sweep of higher arrival rates,
    3000 requests each (seed 6), stopping at the first failed request: where
    served Mrays/s stops following the offered load is the saturation knee.
    Last, the same requests at DEVICE_SWEEP_RATES with rays and results in
    device memory (RT_DEVICE_PTRS), which removes the host copies that bound
    the host sweep."""
    from bvhraytracer_amd import workloads as wl

    t0 = time.time()
    world = [wl.serving_scene(i, args.branching, args.leaf) for i in range(wl.SERVE_SCENES)]
    log(f"serving workload: 16 scenes, {int(wl.serving_sizes().sum())} tris, setup {time.time() - t0:.1f}s")
    srv = rt.Server(hbm_cap_bytes=8 << 30, slots=3)
    sids = [srv.add_scene(w[2]) for w in world]
    # warm the launch path (graph capture) outside the measured window
    for i in range(8):
        rid, _ = srv.submit(sids[i], world[i][3][:4096])
        srv.wait(rid, 30)
    base = serve_run(rt, srv, sids, world, 200, args.serve_seconds, 5)
    log(f"serving at 200 req/s: {base['completed']} requests, {base['mrays_per_s']:.1f} Mrays/s, "
        f"p50 {base['p50_ms']:.2f} ms, p99 {base['p99_ms']:.2f} ms, {base['requests_per_launch']:.2f} requests/launch")
    sweep = []
    if not args.no_serve_sweep:
        for rate in SWEEP_RATES:
            r = serve_run(rt, srv, sids, world, rate, SWEEP_REQUESTS / rate, 6)
            sweep.append(r)
            log(f"serving at {rate} req/s (submitted {r['submitted_rate']}): offered {r['offered_mrays_per_s']} "
                f"Mrays/s, served {r['mrays_per_s']} Mrays/s, p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms, "
                f"{r['requests_per_launch']} requests/launch, failed {r['failed']}")
            if r["failed"]:
                break
    device_sweep = []
    if not args.no_serve_sweep:
        # the same requests with rays and results in device memory (a renderer
        # that spawns rays on the GPU): no host copies on the request path
        import torch
        dev = [(torch.from_numpy(np.ascontiguousarray(w[3]).view(np.float32).reshape(-1, 10)).cuda(),
                torch.from_numpy(np.ascontiguousarray(w[4]).view(np.float32).reshape(-1, 10)).cuda()) for w in world]
        for rate in DEVICE_SWEEP_RATES:
            r = serve_run(rt, srv, sids, world, rate, SWEEP_REQUESTS / rate, 6, device=dev)
            device_sweep.append(r)
            log(f"serving device requests at {rate} req/s (submitted {r['submitted_rate']}): offered "
                f"{r['offered_mrays_per_s']} Mrays/s, served {r['mrays_per_s']} Mrays/s, p50 {r['p50_ms']} ms, "
                f"p99 {r['p99_ms']} ms, {r['requests_per_launch']} requests/launch, failed {r['failed']}")
            if r["failed"]:
                break
        del dev
    srv.close()
    for w in world:
        w[2].close()
    out = {"scenes": 16, "seconds": args.serve_seconds, **base, "sweep": sweep, "device_sweep": device_sweep}
    if device_sweep:
        out["peak_served_device_mrays_per_s"] = max(r["mrays_per_s"] for r in device_sweep)
    if sweep:
        # knee: the first rate whose served throughput falls below 90% of what was offered
        knee = next((r["arrival_rate"] for r in sweep if r["mrays_per_s"] < 0.9 * r["offered_mrays_per_s"]), None)
        out["saturation_knee_req_per_s"] = knee
        out["peak_served_mrays_per_s"] = max(r["mrays_per_s"] for r in sweep)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=8_000_000)
    ap.add_argument("--branching", type=int, default=8)
    ap.add_argument("--leaf", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=2.0, help="target seconds per CPU repetition")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-frame", action="store_true", help="skip the bounce-frame leg")
    ap.add_argument("--no-batch-autotune", action="store_true",
                    help="skip the batched leg with N and L autotuned per scene (~40 s of ingest)")
    ap.add_argument("--no-serve-sweep", action="store_true", help="skip the serving arrival-rate sweep")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards rays "
              "across GPUs by replicating the BVH, a layout not launched here.", file=sys.stderr)
        sys.exit(2)

    import torch

    import bvhraytracer_amd as rt

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    N, L = args.branching, args.leaf
    t0 = time.time()
    tris = rt.interior_scene(2, 1_000_000)
    bvh = rt.Bvh(tris, N, L)
    log(f"scene {len(tris)} tris, N{N}L{L}: {bvh.node_count} nodes, build {time.time() - t0:.1f}s")
    scene = rt.Scene(bvh)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 1920, 1080, 16)
    phits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, phits)
    rays = rt.spawn_diffuse(prim, nrm, pos, phits, args.rays, seed=3)
    n = len(rays)
    log(f"{len(prim)} primary rays ({(phits['slot'] >= 0).mean():.3f} hit), {n} diffuse rays")

    d_rays = torch.from_numpy(rays.view(np.float32).reshape(n, 10)).to(dev)
    d_hits = torch.empty((n, 6), dtype=torch.int32, device=dev)
    d_cnt = torch.empty((n, 4), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    # counts pass: per-ray work for the bytes-per-ray roofline
    scene.intersect_device(d_rays, d_hits, counts=d_cnt, stream=stream)
    counts = d_cnt.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)
    alg_bytes = algorithmic_bytes(counts, bvh.node_memory)
    q_bytes = layout_bytes(counts, bvh.node_memory)
    hits_frac = float((d_hits[:, 2] >= 0).float().mean().item())

    for _ in range(args.warmup):
        scene.intersect_device(d_rays, d_hits, stream=stream)
    torch.cuda.synchronize()
    kms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        scene.intersect_device(d_rays, d_hits, stream=stream)
        kms.append(scene.stats()["last_kernel_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1000.0 / args.steps
    value = n / (ms_per_step / 1000.0) / 1e6
    k_ms = float(np.mean(kms))
    achieved = alg_bytes / (k_ms / 1000.0) / 1e9
    traffic = measured_traffic((N, L, n))
    st = scene.stats()
    log(f"kernel {k_ms:.3f} ms avg, step {ms_per_step:.3f} ms, {value:.1f} Mrays/s, "
        f"{achieved:.0f} GB/s algorithmic, errors {st['error_rays']}")
    # useful bytes: the real child boxes and triangles the same rays test
    # (rt_trace_histograms totals), without the padded slots of k
    _, hist = scene.histograms(rays)
    useful = useful_bytes(n, hist, bvh.node_memory)

    # primary rays of the same scene (coherent), for reference: with the
    # exact-box records (the camera-ray configuration, as rt_render_ambient
    # traces them) and with the quantized records
    d_prim = torch.from_numpy(prim.view(np.float32).reshape(-1, 10)).to(dev)
    d_phit = torch.empty((len(prim), 6), dtype=torch.int32, device=dev)

    def primary_rate(exact):
        scene.intersect_device(d_prim, d_phit, stream=stream, exact_boxes=exact)
        pms = []
        for _ in range(3):
            scene.intersect_device(d_prim, d_phit, stream=stream, exact_boxes=exact)
            pms.append(scene.stats()["last_kernel_ms"])
        return len(prim) / (float(np.median(pms)) / 1000.0) / 1e6

    primary_mrays = primary_rate(True)
    primary_q_mrays = primary_rate(False)

    # every traversal of the boundary on the same scene (kernel time by HIP
    # events, median of 3 after a warm-up): save distance, ray stream and
    # cooperative on the 8M diffuse rays, any hit on the primaries' shadow rays
    def kernel_mrays(d_in, d_out, count, **kw):
        scene.intersect_device(d_in, d_out, stream=stream, **kw)
        ks = []
        for _ in range(3):
            scene.intersect_device(d_in, d_out, stream=stream, **kw)
            ks.append(scene.stats()["last_kernel_ms"])
        return round(count / (float(np.median(ks)) / 1000.0) / 1e6, 1)

    variants = {"closest_per_ray": round(n / (k_ms / 1000.0) / 1e6, 1),
                "closest_exact_boxes": kernel_mrays(d_rays, d_hits, n, exact_boxes=True),
                "save_distance": kernel_mrays(d_rays, d_hits, n, mode="save_distance"),
                "closest_stream": kernel_mrays(d_rays, d_hits, n, stream_sched=True),
                "closest_cooperative": kernel_mrays(d_rays, d_hits, n, coop=True)}
    shadow, _ = rt.spawn_shadow(nrm, pos, phits, (0.0, 9.5, 0.0), 1.0)
    d_sh = torch.from_numpy(shadow.view(np.float32).reshape(-1, 10)).to(dev)
    d_occ = torch.empty(len(shadow), dtype=torch.uint8, device=dev)
    variants["shadow_any_hit"] = kernel_mrays(d_sh, d_occ, len(shadow), mode="any")
    variants["shadow_any_hit_exact_boxes"] = kernel_mrays(d_sh, d_occ, len(shadow), mode="any", exact_boxes=True)
    variants["shadow_any_hit_stream"] = kernel_mrays(d_sh, d_occ, len(shadow), mode="any", stream_sched=True)
    log(f"traversal variants (Mrays/s, kernel time): {variants}")
    del d_sh, d_occ

    # the same 8M rays from host memory (rt_trace without RT_DEVICE_PTRS: rays
    # copied in and hits copied out over PCIe in the call); never `value`
    scene.intersect(rays)
    hts = []
    for _ in range(3):
        t0 = time.perf_counter()
        scene.intersect(rays)
        hts.append(time.perf_counter() - t0)
    host_ptr_mrays = n / float(np.median(hts)) / 1e6
    log(f"host-pointer rt_trace (PCIe in and out): {host_ptr_mrays:.1f} Mrays/s")

    frame = None if args.no_frame else bench_frame(rt, torch, scene, bvh, n, log)
    render_frame = None if args.no_frame else bench_render_frame(rt, scene, bvh, tris, log)
    copy = bench_copy(torch, log)
    batched = None if args.no_batch else bench_batched(rt, torch, args, log)
    serving = None if args.no_serve else bench_serving(rt, args, log)

    cpu = None
    if not args.no_cpu:
        import oracle

        def cpu_time(o, sample, reps, **kw):
            times = []
            for _ in range(reps):
                t = time.perf_counter()
                r = o.trace(rays[:sample], counts=False, **kw)
                times.append(time.perf_counter() - t)
            return sample / float(np.median(times)) / 1e6, int(r["threads"])

        # the traversal with 4-wide SSE node and leaf tests (the reference's
        # ISPC kernel shape, sse4-i32x4) and the scalar restatement, both on
        # every host thread with OpenMP dynamic(4) over 256-ray tiles
        simd, scalar = oracle.Oracle.from_bvh(bvh, simd=True), oracle.Oracle.from_bvh(bvh)
        probe = rays[: min(n, 65536)]
        t = time.perf_counter()
        simd.trace(probe, counts=False)
        rate = len(probe) / max(time.perf_counter() - t, 1e-6)
        sample = int(min(n, max(65536, rate * args.cpu_seconds)))
        cpu_rate, threads = cpu_time(simd, sample, 5)
        scalar_rate, _ = cpu_time(scalar, sample, 5)
        # the stream form (intersectWideAlternative per 256-ray tile) beside it
        stream_rate, _ = cpu_time(simd, sample, 3, stream=True)
        cpu = {"value": round(cpu_rate, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "simd": "4-wide SSE node and leaf tests (gang 4, as the reference's sse4-i32x4 ISPC kernels)",
               "scalar_value": round(scalar_rate, 3), "stream_value": round(stream_rate, 3),
               "sample": f"first {sample} of the {n} diffuse rays, closest-hit per-ray traversal, "
                         f"1 warm-up + 5 timed, median (scalar_value: the scalar restatement, same protocol; "
                         f"stream_value: the SIMD stream traversal per 256-ray tile, 3 timed, median); "
                         f"{cpu_model()}"}
        log(f"cpu baseline {cpu_rate:.3f} Mrays/s per-ray SIMD, {scalar_rate:.3f} scalar, {stream_rate:.3f} "
            f"stream, on {threads} threads ({sample} rays)")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"interior 1M tris, N{N}L{L} G4, {n} cosine-diffuse closest-hit rays "
                               f"(seed 3) from 1920x1080 primary hits, plain per-ray kernel",
                   "rays": n, "triangles": int(len(tris)), "branching": N, "leaf_size": L,
                   "hit_fraction": round(hits_frac, 4),
                   "primary_mrays_1920x1080": round(primary_mrays, 2),
                   "primary_quantized_mrays_1920x1080": round(primary_q_mrays, 2),
                   "host_pointer_mrays": round(host_ptr_mrays, 2), "kernel_ms": round(k_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "algorithmic_bytes": int(alg_bytes),
                     "useful_bytes": int(useful), "useful_frac": round(useful / (k_ms / 1000.0) / 1e9 / HBM_PEAK_GBPS, 4),
                     # what the quantized records move for the same visits
                     "layout_bytes": int(q_bytes), "layout_frac": round(q_bytes / (k_ms / 1000.0) / 1e9 / HBM_PEAK_GBPS, 4),
                     "hbm_frac": hbm_frac(traffic, k_ms),
                     "limiter": limiter_text(MEASURED_COUNTERS.get((N, L, n))),
                     "gather": gather_roofline(MEASURED_COUNTERS.get((N, L, n)), n, k_ms),
                     "counters": MEASURED_COUNTERS.get((N, L, n)),
                     # bytes the L1s read from L2 per step (the counters' pass) over this run's kernel time
                     "l2_read_gbps": (round(MEASURED_COUNTERS[(N, L, n)]["l2_read_bytes"] / (k_ms / 1000.0) / 1e9, 1)
                                      if (N, L, n) in MEASURED_COUNTERS else None)},
        "cpu_baseline": cpu,
        "traversals_mrays": variants,
        "bounce_frame": frame,
        "render_frame": render_frame,
        "hbm_copy": copy,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
