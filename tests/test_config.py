"""config.txt semantics of RayTracer::readConfig (rayTracer.cpp:756-877,
configUtil.h:6-58) and the N/L sweep (rayTracer.cpp:104-122,568-576)."""
import os

from bvhraytracer_amd import raytracer

CONFIG = """#Config file for raytracer.
xRes = 1920
yRes = 1088
minLeafSize = 4
maxLeafSize = 16
leafStep = 4
minBranch = 4
maxBranch = 16
branchStep = 4
workGroupSize = 16
#doPerformanceTest = true
doPerformanceTest = false
doLeafMemoryTests = false
doNodeMemoryTests = true
cacheSize = 512, 256, 12
ambientSampleCount = 1
sortEachSplit = true
leafSplitOption = 1
wideRender = maybe
scenario = 4,8, 9
"""


def test_reference_key_semantics(tmp_path):
    p = tmp_path / "config.txt"
    p.write_text(CONFIG)
    cfg = raytracer.read_config(str(p))
    assert (cfg["xRes"], cfg["yRes"]) == (1920, 1088)
    assert (cfg["minLeafSize"], cfg["maxLeafSize"], cfg["leafStep"]) == (4, 16, 4)
    assert cfg["doPerformanceTest"] is False  # commented line ignored
    # substring match: the file's "doNodeMemoryTests" sets doNodeMemoryTest
    assert cfg["doNodeMemoryTest"] is True and cfg["doLeafMemoryTest"] is False
    assert cfg["sortEachSplit"] is True and cfg["leafSplitOption"] == 1
    assert cfg["wideRender"] is False and cfg["_warnings"]  # neither true nor false
    assert cfg["scenarios"] == [4, 8, 9]
    assert cfg["cacheSizes"] == []  # only read when doCacheAnalysis


def test_shipped_config_values(tmp_path):
    # the sweep values of the reference's shipped config.txt
    p = tmp_path / "config.txt"
    p.write_text("minLeafSize = 4\nmaxLeafSize = 16\nleafStep = 4\nminBranch = 4\nmaxBranch = 16\n"
                 "branchStep = 4\nworkGroupSize = 16\n")
    cfg = raytracer.sanity_clamp(raytracer.read_config(str(p)))
    pairs = list(raytracer.sweep(cfg))
    assert len(pairs) == 16 and pairs[0] == (4, 4) and pairs[1] == (4, 8) and pairs[-1] == (16, 16)


def test_sanity_clamp_swaps_bounds():
    cfg = raytracer.sanity_clamp({"minBranch": 8, "maxBranch": 4, "minLeafSize": 16, "maxLeafSize": 4})
    assert (cfg["minBranch"], cfg["maxBranch"]) == (8, 8)
    assert (cfg["minLeafSize"], cfg["maxLeafSize"]) == (16, 16)


def test_perf_file_fields(tmp_path):
    p = os.path.join(tmp_path, "x_Perf.txt")
    raytracer.write_perf(p, "spheres", "_b4_l4", 4, 4, 4, 4, 4, 1.5, 1.0, 0.25, 8.0, 100002)
    lines = open(p).read().splitlines()
    assert lines[0] == "scenario spheres with branching factor of 4 and leafsize of 4"
    assert lines[2] == "Raytracer total time: 1.5" and lines[-1] == "Triangle Count: 100002"
