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
# per-dispatch values are in profiles/k_trace_N8L8_diffuse8M_hbm_traffic.md.
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_TRAFFIC_KIB = {(8, 8, 8_000_000): {"fetch": 8383360, "write": 397396}}
METRIC = "Mrays/s per MI355X (closest-hit, N8L8) batched vs plain; p99 request latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def header_bytes(M):
    # inner-node header block the kernel loads per visit (rt_layout.h)
    return 32 if M <= 8 else 48


def algorithmic_bytes(counts, M):
    """Bytes the traversal must move per the device layout: ray in (40 B),
    hit out (24 B), 24*k box bytes + header per inner visit, 36*k per leaf."""
    c = counts
    return (64 * len(c) + 24 * c["node_slots"].astype(np.float64).sum()
            + header_bytes(M) * c["nodes"].astype(np.float64).sum()
            + 36 * c["leaf_slots"].astype(np.float64).sum())


def clamp_lognormal(rng, mu, sigma, lo, hi, size):
    return np.clip(np.exp(rng.normal(mu, sigma, size)), lo, hi).astype(np.int64)


def random_camera_rays(rt, rng, n):
    """Primary rays of a random camera inside the 30x10x30 room: n rays of a
    256-wide image (tile order), the last tile row cut to n."""
    pos = rng.uniform([-12.0, 1.0, -12.0], [12.0, 9.0, 12.0]).astype(np.float32)
    d = rng.normal(size=3)
    d[1] *= 0.3
    tgt = (pos + 5.0 * d / np.linalg.norm(d)).astype(np.float32)
    h = max(16, int(np.ceil(n / 256 / 16)) * 16)
    return rt.camera_rays(tuple(pos), tuple(tgt), 256, h, 16)[:n]


def bench_batched(rt, torch, args, log):
    """64 scenes (seeds 100..163), triangles clamp(LogNormal(ln 2e5, 1), 1e4, 2e6),
    rays per scene clamp(LogNormal(ln 6.4e4, 1.2), 4096, 2^20), half primary
    (random camera in the room), half diffuse bounces from those hits.
    plain = one rt_trace per scene; batched = one rt_batch_trace (graph)."""
    rng = np.random.default_rng(4)
    ntris = clamp_lognormal(rng, np.log(2e5), 1.0, 1e4, 2e6, 64)
    nrays = clamp_lognormal(rng, np.log(6.4e4), 1.2, 4096, 1 << 20, 64)
    t0 = time.time()
    scenes, host_rays = [], []
    for i in range(64):
        tris = rt.interior_scene(100 + i, int(ntris[i]))
        bvh = rt.Bvh(tris, args.branching, args.leaf)
        sc = rt.Scene(bvh)
        npri = int(nrays[i]) // 2
        prim = random_camera_rays(rt, np.random.default_rng(1000 + i), npri)
        ph = sc.intersect(prim)
        nrm, pos = bvh.surface(prim, ph)
        if (ph["slot"] >= 0).any():
            diff = rt.spawn_diffuse(prim, nrm, pos, ph, int(nrays[i]) - npri, seed=100 + i)
            rays = np.concatenate([prim, diff])
        else:
            rays = prim
        d_r = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
        d_h = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
        scenes.append((bvh, sc, d_r, d_h))
        host_rays.append((tris, rays))
    total = sum(x[2].shape[0] for x in scenes)
    log(f"batched workload: 64 scenes, {int(ntris.sum())} tris, {total} rays, setup {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()

    def plain():
        for bvh, sc, d_r, d_h in scenes:
            sc.intersect_device(d_r, d_h, stream=stream)

    batch = rt.Batch(64)
    outs = [torch.empty_like(x[3]) for x in scenes]
    items = [(sc, d_r, o, "closest") for (bvh, sc, d_r, d_h), o in zip(scenes, outs)]

    def batched():
        batch.trace(items, graph=True, stream=stream)

    res = {}
    for name, fn in (("plain", plain), ("batched", batched)):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(3, args.steps // 2)):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        res[name] = total / float(np.median(ts)) / 1e6
    same = all(torch.equal(o, x[3]) for o, x in zip(outs, scenes))
    log(f"batched {res['batched']:.1f} Mrays/s vs plain {res['plain']:.1f} Mrays/s, identical={same}")
    batch.close()
    out = {"scenes": 64, "triangles": int(ntris.sum()), "rays": int(total),
           "plain_mrays": round(res["plain"], 2), "batched_mrays": round(res["batched"], 2),
           "speedup": round(res["batched"] / res["plain"], 3), "identical_to_plain": bool(same),
           "n_l": f"N{args.branching}L{args.leaf}", "graph": True}
    if args.batch_autotune:
        out["autotuned"] = batched_autotuned(rt, torch, host_rays, scenes, outs, stream, total, args, log)
    if not args.no_cpu:
        # the CPU baseline of this workload: the oracle per scene, summed
        # (one warm-up pass, then one timed pass over all 64 scenes)
        import oracle

        orcs = [(oracle.Oracle.from_bvh(bvh), rays) for (bvh, _, _, _), (_, rays) in zip(scenes, host_rays)]
        for o, rays in orcs:
            o.trace(rays, counts=False)
        t = time.perf_counter()
        for o, rays in orcs:
            used = o.trace(rays, counts=False)["threads"]
        cpu_s = time.perf_counter() - t
        out["cpu_port_mrays"] = round(total / cpu_s / 1e6, 3)
        out["cpu_threads"] = int(used)
        log(f"batched workload on the CPU oracle: {out['cpu_port_mrays']} Mrays/s on {used} threads")
    for bvh, sc, d_r, d_h in scenes:
        sc.close()
    return out


def batched_autotuned(rt, torch, host_rays, scenes, ref_outs, stream, total, args, log):
    """The same 64 scenes with N and L chosen per scene by the autotuner
    (candidates N4L4, N4L8, N8L4, N8L8 timed on 16384 of the scene's own rays),
    traced in one batched graph launch that mixes the chosen node/leaf
    memories.  Tuning and the rebuilds are ingest cost, outside the timing."""
    from bvhraytracer_amd.autotune import Autotuner

    t0 = time.time()
    tuner = Autotuner(nl=((4, 4), (4, 8), (8, 4), (8, 8)), blocks=(256,), windows=(16,), schedules=(16,),
                      sample=16384, reps=2)
    tuned, choice = [], {}
    for (tris, rays), (bvh0, sc0, d_r, d_h) in zip(host_rays, scenes):
        best = tuner.tune_scene(tris, rays)
        key = f"N{best['N']}L{best['L']}"
        choice[key] = choice.get(key, 0) + 1
        bvh = rt.Bvh(tris, best["N"], best["L"])
        sc = rt.Scene(bvh)
        tuned.append((bvh, sc, d_r, torch.empty_like(d_h)))
    tune_s = time.time() - t0
    batch = rt.Batch(64)
    items = [(sc, d_r, o, "closest") for (bvh, sc, d_r, o) in tuned]
    for _ in range(max(1, args.warmup)):
        batch.trace(items, graph=True, stream=stream)
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(3, args.steps // 2)):
        t = time.perf_counter()
        batch.trace(items, graph=True, stream=stream)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    mrays = total / float(np.median(ts)) / 1e6
    # the hit triangles are the same whatever the tree, up to exact ties
    same_tri = sum(int((o[:, 0] == r[:, 0]).sum().item()) for (b, s, d, o), r in zip(tuned, ref_outs)) / total
    log(f"batched autotuned {mrays:.1f} Mrays/s, choices {choice}, tuning {tune_s:.1f}s")
    batch.close()
    for bvh, sc, d_r, o in tuned:
        sc.close()
    return {"batched_mrays": round(mrays, 2), "choices": choice, "tune_seconds": round(tune_s, 1),
            "same_hit_triangle_fraction": round(same_tri, 6)}


def bench_frame(rt, torch, scene, bvh, n_bounce, log, reps=3):
    """One bounce frame: 1920x1080 primaries -> closest hit -> hit surfaces ->
    n_bounce diffuse rays -> closest hit.  device: every stage on the GPU on
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
            scene.intersect_device(d_r, d_h, stream=stream, sync=False)
            d_n, d_p = scene.surface_device(d_r, d_h, stream=stream)
            d_b = rt.spawn_diffuse_device(d_r, d_n, d_p, d_h, n_bounce, seed=3, stream=stream)
            scene.intersect_device(d_b, d_bh, stream=stream, sync=False)
        stream.synchronize()

    def host_frame():
        prim = rt.camera_rays(pos, tgt, W, H, 16)
        hits = scene.intersect(prim)
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


def bench_serving(rt, args, log):
    """16 resident scenes (100k..2M tris log-spaced, seeds 200..215), Poisson
    arrivals at 200 req/s for args.serve_seconds (seed 5), rays per request
    clamp(LogNormal(ln 2.3e4, 1.3), 1024, 2^19), half shadow (any-hit), half
    closest-hit; latency from submit to results in host memory."""
    import threading
    t0 = time.time()
    counts = np.round(np.exp(np.linspace(np.log(1e5), np.log(2e6), 16))).astype(np.int64)
    world = []
    for i in range(16):
        tris = rt.interior_scene(200 + i, int(counts[i]))
        bvh = rt.Bvh(tris, args.branching, args.leaf)
        sc = rt.Scene(bvh)
        prim = random_camera_rays(rt, np.random.default_rng(2000 + i), 1 << 19)
        ph = sc.intersect(prim)
        nrm, pos = bvh.surface(prim, ph)
        diff = rt.spawn_diffuse(prim, nrm, pos, ph, 1 << 19, seed=200 + i)
        shadow, _ = rt.spawn_shadow(nrm, pos, ph, (0.0, 9.5, 0.0), 1.0)
        shadow = np.resize(shadow, 1 << 19)
        world.append((bvh, sc, np.concatenate([prim, diff]), shadow))
    log(f"serving workload: 16 scenes, {int(counts.sum())} tris, setup {time.time() - t0:.1f}s")
    rng = np.random.default_rng(5)
    n_req = int(200 * args.serve_seconds)
    gaps = rng.exponential(1 / 200.0, n_req)
    sizes = clamp_lognormal(rng, np.log(2.3e4), 1.3, 1024, 1 << 19, n_req)
    scene_of = rng.integers(0, 16, n_req)
    is_shadow = rng.random(n_req) < 0.5
    offs = [int(rng.integers(0, (1 << 20) - sizes[i] + 1)) if not is_shadow[i]
            else int(rng.integers(0, (1 << 19) - sizes[i] + 1)) for i in range(n_req)]
    srv = rt.Server(hbm_cap_bytes=8 << 30, slots=3)
    sids = [srv.add_scene(w[1]) for w in world]
    # warm the launch path (graph capture) outside the measured window
    for i in range(8):
        rid, _ = srv.submit(sids[i], world[i][2][:4096])
        srv.wait(rid, 30)
    srv.stats(reset=True)
    done_q = []
    lock = threading.Lock()

    def waiter():
        k = 0
        while k < n_req:
            with lock:
                ready = k < len(done_q)
                rid = done_q[k] if ready else None
            if not ready:
                time.sleep(0.0005)
                continue
            srv.wait(rid, 120)
            k += 1

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    at = start
    for i in range(n_req):
        at += gaps[i]
        dt = at - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        w = world[scene_of[i]]
        src = w[3] if is_shadow[i] else w[2]
        rid, _ = srv.submit(sids[scene_of[i]], src[offs[i]:offs[i] + sizes[i]],
                            "any" if is_shadow[i] else "closest")
        with lock:
            done_q.append(rid)
    th.join()
    st = srv.stats()
    srv.close()
    for w in world:
        w[1].close()
    log(f"serving: {st['completed']} requests, {st['mrays_per_s']:.1f} Mrays/s, p50 {st['p50_ms']:.2f} ms, "
        f"p99 {st['p99_ms']:.2f} ms, {st['mean_requests_per_launch']:.1f} requests/launch")
    return {"scenes": 16, "requests": int(st["completed"]), "arrival_rate": 200,
            "seconds": args.serve_seconds, "offered_mrays_per_s": round(float(sizes.sum()) / args.serve_seconds / 1e6, 2),
            "mrays_per_s": round(st["mrays_per_s"], 2), "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
            "requests_per_launch": round(st["mean_requests_per_launch"], 2), "failed": int(st["failed"])}


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
    ap.add_argument("--batch-autotune", action="store_true",
                    help="also run the batched leg with N and L autotuned per scene (adds ~1 min of ingest)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
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
    mt = MEASURED_TRAFFIC_KIB.get((N, L, n))
    traffic = None if mt is None else int((mt["fetch"] * FETCH_SCALE + mt["write"] * WRITE_SCALE) * 1024)
    st = scene.stats()
    log(f"kernel {k_ms:.3f} ms avg, step {ms_per_step:.3f} ms, {value:.1f} Mrays/s, "
        f"{achieved:.0f} GB/s algorithmic, errors {st['error_rays']}")

    # primary rays of the same scene (coherent), for reference
    d_prim = torch.from_numpy(prim.view(np.float32).reshape(-1, 10)).to(dev)
    d_phit = torch.empty((len(prim), 6), dtype=torch.int32, device=dev)
    scene.intersect_device(d_prim, d_phit, stream=stream)
    pms = []
    for _ in range(3):
        scene.intersect_device(d_prim, d_phit, stream=stream)
        pms.append(scene.stats()["last_kernel_ms"])
    primary_mrays = len(prim) / (float(np.median(pms)) / 1000.0) / 1e6

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
                "save_distance": kernel_mrays(d_rays, d_hits, n, mode="save_distance"),
                "closest_stream": kernel_mrays(d_rays, d_hits, n, stream_sched=True),
                "closest_cooperative": kernel_mrays(d_rays, d_hits, n, coop=True)}
    shadow, _ = rt.spawn_shadow(nrm, pos, phits, (0.0, 9.5, 0.0), 1.0)
    d_sh = torch.from_numpy(shadow.view(np.float32).reshape(-1, 10)).to(dev)
    d_occ = torch.empty(len(shadow), dtype=torch.uint8, device=dev)
    variants["shadow_any_hit"] = kernel_mrays(d_sh, d_occ, len(shadow), mode="any")
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

        o = oracle.Oracle.from_bvh(bvh)
        probe = rays[: min(n, 65536)]
        t = time.perf_counter()
        r = o.trace(probe, counts=False)
        rate = len(probe) / max(time.perf_counter() - t, 1e-6)
        sample = int(min(n, max(65536, rate * args.cpu_seconds)))
        times = []
        for _ in range(5):
            t = time.perf_counter()
            r = o.trace(rays[:sample], counts=False)
            times.append(time.perf_counter() - t)
        cpu_rate = sample / float(np.median(times)) / 1e6
        # the stream form (intersectWideAlternative per 256-ray tile) beside it
        st = []
        for _ in range(3):
            t = time.perf_counter()
            o.trace(rays[:sample], stream=True, counts=False)
            st.append(time.perf_counter() - t)
        stream_rate = sample / float(np.median(st)) / 1e6
        cpu = {"value": round(cpu_rate, 3), "unit": "Mrays/s", "cores": int(r["threads"]),
               "kind": "port", "stream_value": round(stream_rate, 3),
               "sample": f"first {sample} of the {n} diffuse rays, closest-hit per-ray traversal, "
                         f"1 warm-up + 5 timed, median (stream_value: the stream traversal per "
                         f"256-ray tile, 3 timed, median); {cpu_model()}"}
        log(f"cpu baseline {cpu_rate:.3f} Mrays/s per-ray, {stream_rate:.3f} stream, on {r['threads']} threads "
            f"({sample} rays)")

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
                   "host_pointer_mrays": round(host_ptr_mrays, 2), "kernel_ms": round(k_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic},
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
