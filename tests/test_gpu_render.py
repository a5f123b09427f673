"""The CameraFast render driver end to end on the GPU (cameraFast.h:170-582,
829-852): primary closest-hit rays, hit surfaces, ambient any-hit rays and
the saved _Perf.png, against the oracle tracing the same rays; and the
RayTracer sweep writing _Perf.txt and _Perf.png for each (N, L)."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import oracle_for
from tests.test_image import decode_png

pytestmark = pytest.mark.gpu

POS, TGT = (0.0, 4.0, 32.0), (0.0, 0.0, 0.0)


def test_render_image_png_matches_oracle(rt, tmp_path):
    from bvhraytracer_amd.raytracer import CameraFast, ambient_image, tile_order_to_rows

    W, H, wg, samples, dist = 160, 120, 16, 3, 4.0
    print(f"render {W}x{H} wg {wg} samples {samples}", flush=True)
    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 4, 4)
    scene = rt.Scene(bvh)
    cam = CameraFast([POS], [TGT], W, H, wg)
    cam.render_images(scene, bvh, samples, dist, samples=1, save_image=True, path=str(tmp_path), name="spheres")
    png = tmp_path / "spheres_c0_Perf.png"
    got = decode_png(png.read_bytes())

    orc = oracle_for(bvh)
    rays = rt.camera_rays(POS, TGT, W, H, wg)
    oh = orc.trace(rays)["hits"]
    ref = np.zeros(len(rays), rt.HIT_DTYPE)
    for f in ("slot", "leaf", "t", "b1", "b2"):
        ref[f] = oh[f]
    hit = ref["slot"] >= 0
    assert 0 < hit.sum() < len(hit)
    nrm, pos = bvh.surface(rays, ref)
    occl = np.zeros(len(rays), np.uint16)
    for a in range(samples):
        amb, src = rt.spawn_ambient(nrm, pos, ref, a, dist)
        assert np.array_equal(src, np.nonzero(hit)[0])
        occl[src] += orc.trace(amb, mode=oracle.ORC_ANY)["occluded"].astype(np.uint16)
    assert 0 < occl.sum() < samples * hit.sum()
    want = tile_order_to_rows(ambient_image(occl, hit, samples), W, H, wg)
    assert np.array_equal(got[..., 0], want)
    assert np.array_equal(got[..., 1], want) and np.array_equal(got[..., 2], want)
    assert np.all(got[..., 3] == 255)
    scene.close()


def test_raytracer_sweep_writes_reports(rt, tmp_path):
    from bvhraytracer_amd.raytracer import RayTracer

    cfg = {"minLeafSize": 4, "maxLeafSize": 8, "leafStep": 4, "minBranch": 4, "maxBranch": 4, "branchStep": 4,
           "xRes": 64, "yRes": 48, "workGroupSize": 16, "ambientSampleCount": 1, "saveImage": True}
    out = tmp_path / "perf"
    res = RayTracer(cfg, out_dir=str(out)).run(("spheres",))
    assert set(res) == {("spheres", 4, 4), ("spheres", 4, 8)}
    files = sorted(os.listdir(out))
    assert "spheres_c0_Perf.png" in files
    perf = sorted(os.listdir(out / "WorkGroupSize_16_Normal"))
    assert perf == ["spheres_b4_l4_mb4_ml4_Perf.txt", "spheres_b4_l8_mb4_ml8_Perf.txt"]
    img = decode_png((out / "spheres_c0_Perf.png").read_bytes())
    assert img.shape == (48, 64, 4) and img[..., 0].max() > 0


def test_workgroup_analysis_end_to_end(rt, tmp_path):
    """workgroup_analysis on a 64x48 frame in 16x16 tiles with two ambient
    samples: the primary records equal the oracle's, each tile's secondary
    records are the two samples' records back to back, and the four files
    hold one line per tile (whiskers, unique work) or per step."""
    from bvhraytracer_amd.raytracer import workgroup_analysis

    W, H, wg = 64, 48, 16
    print("workgroup analysis 64x48 wg 16", flush=True)
    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 8, 8)
    scene = rt.Scene(bvh)
    prim, sec, files = workgroup_analysis(scene, bvh, POS, TGT, W, H, wg, 2, 4.0, str(tmp_path), "spheres",
                                          "_b8_l8")
    tiles = (W // wg) * (H // wg)
    rays = rt.camera_rays(POS, TGT, W, H, wg)
    ids, lens = oracle_for(bvh).visit_sequences(rays)
    want = oracle.tile_step_records(ids, lens, wg * wg, np.ones(len(rays), bool))
    assert len(prim) == len(sec) == tiles
    for g, w in zip(prim, want):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))
    for t in sec:
        assert int(t["terminations"].sum()) == 2 * wg * wg
    lines = [open(f).read().splitlines() for f in files]
    assert len(lines[0]) == len(lines[1]) == len(lines[3]) == tiles + 1
    assert len(lines[2]) == max(len(t) for t in prim) + 1
    scene.close()


@pytest.mark.parametrize("mode,wide,samples", [("closest", False, 3), ("save_distance", False, 1), ("closest", True, 2),
                                               ("closest", False, 0)])
def test_device_frame_equals_host_pipeline(rt, mode, wide, samples):
    """rt_render_ambient (camera rays, trace, surfaces, ambient rays, trace,
    gray values, all on the GPU) gives the host-driven frame's gray values
    exactly, for a partial last tile row too."""
    from bvhraytracer_amd.raytracer import CameraFast

    print(f"device frame {mode} wide={wide} samples={samples}", flush=True)
    tris = rt.interior_scene(3, 60_000, 40)
    bvh = rt.Bvh(tris, 8, 4)
    scene = rt.Scene(bvh)
    cam = CameraFast([(0.0, 5.0, -14.0)], [(0.0, 3.0, 0.0)], 96, 40, 16)
    host = cam.render_image(scene, bvh, 0, samples, 2.0, mode=mode)[3]
    total, primary, _, dev = cam.render_image_device(scene, 0, samples, 2.0, mode=mode, wide=wide)
    assert np.array_equal(dev, host)
    assert 0 < primary <= total
    if samples:
        # some pixels lit, some occluded or missed; with several samples,
        # partial occlusion too
        assert 0 < (host > 0).sum() < len(host)
        assert len(np.unique(host)) > (2 if samples > 1 else 1)
    scene.close()
