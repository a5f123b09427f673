"""The C++ renderer layer (include/rt_renderer.hpp, the bvhrt program) reads
config.txt exactly as the Python driver does; both follow
RayTracer::readConfig (rayTracer.cpp:756-877, configUtil.h:6-58)."""
import os
import subprocess

import pytest

from bvhraytracer_amd import raytracer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BVHRT = os.path.join(ROOT, "bvhraytracer_amd", "bvhrt")

DEFAULTS = dict(minLeafSize=4, maxLeafSize=4, leafStep=4, minBranch=4, maxBranch=4, branchStep=4,
                subdivisionStart=0, subdivisionEnd=0, subdivisionStep=1, renderType=0, ambientSampleCount=1,
                workGroupSize=16, leafSplitOption=1, xRes=1920, yRes=1088, saveImage=False,
                saveDepthDetailedImage=False, bvhAnalysis=False, saveBvhImage=False, sortEachSplit=True,
                renderAnalysisImage=False, castShadows=False, doPerformanceTest=True, doNodeMemoryTest=False,
                doLeafMemoryTest=False, saveRayTimes=False, doWorkGroupAnalysis=False, saveDistance=False,
                wideRender=False, wideAlternative=True, renderAllOptions=False, doCacheAnalysis=False)

CONFIGS = [
    # the reference's key semantics: substring keys, comments, bad booleans
    """#Config file for raytracer.
xRes = 1920
yRes = 1088
minLeafSize = 4
maxLeafSize = 16
leafStep = 4
minBranch = 4
maxBranch = 16
branchStep = 4
#doPerformanceTest = true
doPerformanceTest = false
doLeafMemoryTests = false
doNodeMemoryTests = true
cacheSize = 512, 256, 12
wideRender = maybe
scenario = 4,8, 9
""",
    # cache sizes only with doCacheAnalysis; a size not divisible by 8 is dropped
    """doCacheAnalysis = true
cacheSize = 512, 12, 256
scenario = 0
xRes=640
yRes = 480 # trailing text after the number
saveImage = true
""",
    # two keys in one line both take the value after the first '='
    """minLeafSize maxLeafSize = 8
minBranch=12
maxBranch = 4
ambientSampleCount = 3
doCacheAnalysis = true
cacheSize = 20
""",
]


def cpp_config(path):
    out = subprocess.run([BVHRT, "--print-config", path], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    cfg, warnings = {}, []
    for line in out.stdout.splitlines():
        k, v = line.split("=", 1)
        if k == "warning":
            warnings.append(v)
        elif k in ("scenarios", "cacheSizes"):
            cfg[k] = [int(x) for x in v.split(",")] if v else []
        elif v in ("true", "false"):
            cfg[k] = v == "true"
        else:
            cfg[k] = int(v)
    return cfg, warnings


@pytest.mark.parametrize("i", range(len(CONFIGS)))
def test_cpp_reads_config_like_python(tmp_path, i):
    p = tmp_path / "config.txt"
    p.write_text(CONFIGS[i])
    got, warnings = cpp_config(str(p))
    want = raytracer.read_config(str(p))
    for k, default in DEFAULTS.items():
        assert got[k] == want.get(k, default), k
    assert got["scenarios"] == want["scenarios"]
    assert got["cacheSizes"] == want["cacheSizes"]
    assert warnings == want["_warnings"]


def test_cpp_config_errors(tmp_path):
    out = subprocess.run([BVHRT, "--print-config", str(tmp_path / "missing.txt")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 2 and "cannot open" in out.stderr
    p = tmp_path / "config.txt"
    p.write_text("xRes = wide\n")
    out = subprocess.run([BVHRT, "--print-config", str(p)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "no integer" in out.stderr
