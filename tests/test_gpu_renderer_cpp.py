"""The C++ renderer (bvhrt: RayTracer + CameraFast over the C ABI) on the
GPU: a config.txt performance run writes the reference's _Perf.txt files
and the _Perf.png, and the image equals the Python driver's render of the
same BVH pixel for pixel."""
import os
import subprocess

import numpy as np
import pytest

from tests.test_image import decode_png

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BVHRT = os.path.join(ROOT, "bvhraytracer_amd", "bvhrt")

CONFIG = """xRes = 64
yRes = 48
workGroupSize = 16
minLeafSize = 4
maxLeafSize = 4
leafStep = 4
minBranch = 4
maxBranch = 8
branchStep = 4
ambientSampleCount = 2
saveImage = true
doPerformanceTest = true
scenario = 0
"""


def test_bvhrt_performance_run(rt, tmp_path):
    from bvhraytracer_amd.raytracer import CameraFast, ambient_distance, tile_order_to_rows

    cfg = tmp_path / "config.txt"
    cfg.write_text(CONFIG)
    out = tmp_path / "perf"
    print("bvhrt spheres 64x48 N4,N8 L4", flush=True)
    r = subprocess.run([BVHRT, str(cfg), str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 2 and lines[0].startswith("scenario 0 N4 L4") and lines[1].startswith("scenario 0 N8 L4")

    tris = rt.spheres_scene(0)
    perf_dir = out / "WorkGroupSize_16_Normal"
    assert sorted(os.listdir(perf_dir)) == ["spheres_b4_l4_mb4_ml4_Perf.txt", "spheres_b8_l4_mb8_ml4_Perf.txt"]
    for N in (4, 8):
        bvh = rt.Bvh(tris, N, 4, sah_factor=rt.load_sah_factor(4, N, 4))
        text = (perf_dir / f"spheres_b{N}_l4_mb{N}_ml4_Perf.txt").read_text().splitlines()
        assert text[0] == f"scenario spheres with branching factor of {N} and leafsize of 4"
        assert text[1] == f"with a gang size of 4, a node meory of {N}, and a leaf memory of 4"
        assert text[6] == "Average BVH depth: %g" % np.float32(bvh.average_depth)
        assert text[7] == "Triangle Count: 100002"
        assert float(text[2].split(": ")[1]) > 0

    # the PNG is written per (N, L) under the same name: the last one is N8 L4
    got = decode_png((out / "spheres_c0_Perf.png").read_bytes())
    bvh = rt.Bvh(tris, 8, 4, sah_factor=rt.load_sah_factor(4, 8, 4))
    scene = rt.Scene(bvh)
    cam = CameraFast([(0.0, 4.0, 32.0)], [(0.0, 0.0, 0.0)], 64, 48, 16)
    img = cam.render_image(scene, bvh, 0, 2, ambient_distance(tris))[3]
    want = tile_order_to_rows(img, 64, 48, 16)
    assert 0 < (want > 0).sum() < want.size
    assert np.array_equal(got[..., 0], want) and np.all(got[..., 3] == 255)
    scene.close()


def test_bvhrt_analysis_matches_python(rt, tmp_path):
    """renderAnalysisImage with the wide renderer and the workgroup analysis:
    bvhrt (C++) and the Python RayTracer write byte-identical _Info.txt and
    WorkGroup*.txt files for the same config."""
    from bvhraytracer_amd import raytracer

    cfg_text = ("xRes = 64\nyRes = 48\nworkGroupSize = 16\nminLeafSize = 4\nmaxLeafSize = 4\nleafStep = 4\n"
                "minBranch = 8\nmaxBranch = 8\nbranchStep = 4\nambientSampleCount = 2\ndoPerformanceTest = false\n"
                "renderAnalysisImage = true\nwideRender = true\ndoWorkGroupAnalysis = true\nscenario = 0\n")
    cfg = tmp_path / "config.txt"
    cfg.write_text(cfg_text)
    print("bvhrt analysis spheres 64x48 N8L4", flush=True)
    r = subprocess.run([BVHRT, str(cfg), str(tmp_path / "cpp")], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    raytracer.RayTracer(raytracer.read_config(str(cfg)), out_dir=str(tmp_path / "py")).run(("spheres",))
    names = ["spheres_b8_l4_Info.txt"] + [f"WorkGroupSize_16_Version_1/spheres_b8_l4_c0_{k}.txt" for k in
                                          ("PrimaryWorkGroupWiskerPlot", "SecondaryWorkGroupWiskerPlot",
                                           "WorkGroupData", "WorkGroupUniqueWork")]
    for n in names:
        py = (tmp_path / "py" / n).read_text()
        assert len(py.splitlines()) > 5
        assert (tmp_path / "cpp" / n).read_text() == py, n


def test_bvhrt_memory_tests(rt, tmp_path):
    """doLeafMemoryTest / doNodeMemoryTest: one _Perf.txt per (node memory,
    leaf memory) from the minimum up to 3 gangs more, at most 16."""
    cfg = tmp_path / "config.txt"
    cfg.write_text("xRes = 32\nyRes = 32\nworkGroupSize = 16\nminLeafSize = 8\nmaxLeafSize = 8\nleafStep = 4\n"
                   "minBranch = 12\nmaxBranch = 12\nbranchStep = 4\nambientSampleCount = 1\n"
                   "doPerformanceTest = true\ndoLeafMemoryTest = true\ndoNodeMemoryTest = true\nscenario = 0\n")
    print("bvhrt memory tests N12 L8", flush=True)
    r = subprocess.run([BVHRT, str(cfg), str(tmp_path / "out")], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    files = sorted(os.listdir(tmp_path / "out" / "WorkGroupSize_16_Normal"))
    want = sorted(f"spheres_b12_l8_mb{m}_ml{l}_Perf.txt" for l in (8, 12, 16) for m in (12, 16))
    assert files == want
