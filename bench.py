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
              "across GPUs by replicating the BVH, which this build does not launch.", file=sys.stderr)
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
        cpu = {"value": round(cpu_rate, 3), "unit": "Mrays/s", "cores": int(r["threads"]),
               "kind": "port",
               "sample": f"first {sample} of the {n} diffuse rays, closest-hit per-ray traversal, "
                         f"1 warm-up + 5 timed, median; {cpu_model()}"}
        log(f"cpu baseline {cpu_rate:.3f} Mrays/s on {r['threads']} threads ({sample} rays)")

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
                   "kernel_ms": round(k_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
