"""Builds librt.so (host C++ + HIP for gfx950) and the bvhrt program in-tree.

Host sources are compiled with g++, HIP sources with hipcc, and the shared
library is linked with hipcc; bvhrt (the C++ renderer's command line) links
librt.so from its own directory.  Every float path is built with
-ffp-contract=off: the reference kernels (ISPC, sse4-i32x4) have no FMA, and
bit-exact parity depends on the same operation sequence.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librt.so")
CLI = os.path.join(HERE, "bvhrt")
OBJ = os.path.join(HERE, "_build")

HOST_SOURCES = ["scene.cpp", "builder.cpp", "host_api.cpp", "renderer.cpp", "report_api.cpp"]
HIP_SOURCES = ["rt_kernels.hip", "rt_api.hip", "rt_serve.hip", "rt_shade.hip", "rt_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CXXFLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-pthread",
            "-Wall", "-Wextra", "-Wno-unused-parameter"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-parameter", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, out=OUT, obj_dir=OBJ, defines=()):
    """Compiles and links the library to `out`; `defines` are extra -D
    options for the HIP sources (kernel variants)."""
    os.makedirs(obj_dir, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        jobs.append(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        jobs.append([HIPCC, *HIPFLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for text in ex.map(_run, jobs):
            if verbose and text.strip():
                print(text)
    objs = [os.path.join(obj_dir, s + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,librt.so", "-o", tmp, *objs, "-pthread",
          "-lz"])
    os.replace(tmp, out)
    if out == OUT:
        # the program finds librt.so next to itself ($ORIGIN), never the system librt
        tmp = CLI + ".tmp"
        _run(["g++", *CXXFLAGS, "-o", tmp, os.path.join(CSRC, "bvhrt_main.cpp"), out, "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, CLI)
    return out


def build_oracle():
    """Builds the oracle (test infrastructure) with its own Makefile."""
    _run(["make", "-s", "-C", os.path.join(os.path.dirname(HERE), "oracle")])


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_oracle()
    print(OUT)
