"""The C ABI: librt.so loads and exports every function the C headers in
include/ declare (rt.h, rt_report.h), and the Python records match the C
struct layouts (no GPU call)."""
import os
import re
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(rt_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_all_declared_symbols_exported(rt):
    names = declared_functions()
    assert len(names) >= 25 and "rt_trace" in names and "rt_report_write_info" in names
    lib = rt.lib()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    bound = {s[0] for s in rt._lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_struct_layouts_match_c(rt):
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "rt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(rt_ray), sizeof(rt_hit), sizeof(rt_counts),
         sizeof(rt_bvh_desc), sizeof(rt_stats_t), sizeof(rt_build_params), sizeof(rt_camera));
  printf("%zu %zu %zu %zu\n", offsetof(rt_bvh_desc, tri_float_count), offsetof(rt_bvh_desc, tri_ids),
         offsetof(rt_bvh_desc, max_depth), offsetof(rt_stats_t, error_rays));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    import ctypes
    L = rt._lib
    got = [rt.RAY_DTYPE.itemsize, rt.HIT_DTYPE.itemsize, rt.COUNTS_DTYPE.itemsize,
           ctypes.sizeof(L.RtBvhDesc), ctypes.sizeof(L.RtStats), ctypes.sizeof(L.RtBuildParams),
           ctypes.sizeof(L.RtCamera),
           L.RtBvhDesc.tri_float_count.offset, L.RtBvhDesc.tri_ids.offset, L.RtBvhDesc.max_depth.offset,
           L.RtStats.error_rays.offset]
    assert [int(x) for x in out] == got


def test_errors_are_reported_not_raised_across_the_abi(rt):
    lib = rt.lib()
    assert lib.rt_trace(None, None, 0, 0, None, None, None) == -1
    assert b"null scene" in lib.rt_last_error()
    assert lib.rt_release(None) == 0
    # a layout violation is rejected before any device work
    arr = rt.FastNodeArrays(4, 4, 4, np.zeros((1, 24), np.float32), np.zeros(1, np.uint32),
                            np.zeros((1, 3, 4), np.int8), np.zeros(1, np.uint16),
                            np.zeros(35, np.float32), np.zeros(4, np.int32))
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.rt_upload_bvh(ctypes.byref(arr.desc), ctypes.byref(h))
    assert rc == -5 and b"leaf blocks" in lib.rt_last_error()


def test_cpp_layer_compiles_and_links(tmp_path, rt):
    """include/rt_renderer.hpp compiles on its own and its functions resolve
    in librt.so (config parsing and the sweep run without a GPU)."""
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include <cstdio>
#include <fstream>
#include "rt_renderer.hpp"
int main(int argc, char** argv) {
  { std::ofstream f(argv[1]); f << "minLeafSize = 4\nmaxLeafSize = 8\nminBranch = 4\nmaxBranch = 8\n"; }
  rt::RenderConfig c;
  if (!rt::read_config(argv[1], c)) return 3;
  for (auto& p : rt::sweep(c)) std::printf("%d,%d ", p.first, p.second);
  std::printf("%d %zu\n", rt::ambient_gray(true, 1, 3), sizeof(rt_step_stat));
  try { rt::tile_order_to_rows(std::vector<uint8_t>(5), 4, 1, 2); } catch (const rt::Error&) { std::printf("error\n"); }
  return 0;
}
''')
    exe = tmp_path / "use"
    lib = os.path.join(ROOT, "bvhraytracer_amd", "librt.so")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), lib,
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path / "config.txt")], capture_output=True, text=True, check=True)
    assert out.stdout.split("\n")[:2] == ["4,4 4,8 8,4 8,8 169 20", "error"]


def test_integration_table_names_every_entry_point():
    """INTEGRATION.md's entry-point table names every function include/ declares
    (each with the reference interface it replaces, or "no counterpart")."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    table = text[text.index("## 4. Entry points"):]
    missing = [n for n in declared_functions() if f"`{n}`" not in table]
    assert not missing, missing
