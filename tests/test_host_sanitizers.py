"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (leaks
included): tests/sanitize/host_driver.cpp drives the builder, the scene
generator, the host ray functions and the oracle; renderer_driver.cpp the
host-only parts of the C++ renderer layer.  Both are built with g++ from the
sources (no GPU code) and must run clean.  The host driver takes about six
minutes under the sanitizers, so the test runs when BVHRT_SANITIZE=1 is set
and is skipped otherwise."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bvhraytracer_amd", "csrc")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-ffp-contract=off"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")

pytestmark = pytest.mark.skipif(os.environ.get("BVHRT_SANITIZE") != "1",
                                reason="sanitizer builds take minutes; set BVHRT_SANITIZE=1")


def _run(cmd, **kw):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    assert r.returncode == 0, r.stdout[-4000:]
    return r.stdout


def test_host_code_clean(tmp_path):
    obj = tmp_path / "cpu_ref.o"
    _run(["gcc", *SAN, "-fopenmp", "-c", os.path.join(ROOT, "oracle", "cpu_ref.c"), "-o", str(obj)])
    exe = tmp_path / "host_driver"
    _run(["g++", *SAN, "-std=c++17", "-fopenmp", "-pthread", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(ROOT, "oracle"), "-I", CSRC, os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"),
          *[os.path.join(CSRC, f) for f in ("scene.cpp", "builder.cpp", "host_api.cpp")], str(obj), "-o", str(exe)])
    out = _run([str(exe), "quick"], env=ENV, timeout=1800)
    assert "cases clean" in out


def test_renderer_host_parts_clean(tmp_path, rt):
    exe = tmp_path / "renderer_driver"
    lib = os.path.join(ROOT, "bvhraytracer_amd")
    _run(["g++", *SAN, "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          os.path.join(ROOT, "tests", "sanitize", "renderer_driver.cpp"),
          *[os.path.join(CSRC, f) for f in ("renderer.cpp", "report_api.cpp", "scene.cpp", "builder.cpp",
                                            "host_api.cpp")],
          "-L", lib, "-l:librt.so", f"-Wl,-rpath,{lib}", "-lz", "-o", str(exe)])
    out = _run([str(exe), str(tmp_path)], env=ENV, timeout=600)
    assert "renderer host parts clean" in out
