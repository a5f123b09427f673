"""The workgroup analysis files (CameraData::workGroupAnalysisPerImage,
cameraData.h:660-1109) from hand-made step records of a 4x2 image in 2x2
tiles; every expected line is worked out by hand in the comments."""
import numpy as np

STEP = np.dtype([("node_work", "<u4"), ("node_unique", "<u4"), ("leaf_work", "<u4"),
                 ("leaf_unique", "<u4"), ("terminations", "<u4")])


def recs(*rows):
    return np.array(list(rows), STEP)


def test_workgroup_files_hand_case(tmp_path, rt):
    from bvhraytracer_amd.raytracer import write_workgroup_analysis

    primary = [
        # finished at steps 2, 2, 3, 3
        recs((4, 1, 0, 0, 0), (4, 1, 0, 0, 0), (2, 2, 2, 1, 2), (0, 0, 2, 2, 2)),
        # one ray not started, three finished at step 2
        recs((3, 1, 0, 0, 1), (3, 1, 0, 0, 0), (0, 0, 3, 2, 3)),
    ]
    secondary = [
        recs((0, 1, 0, 0, 4)),  # no ray started
        recs((4, 1, 0, 0, 0), (4, 1, 0, 0, 0), (0, 0, 4, 3, 4)),
    ]
    pv = (np.array([1, 1, 2, 2, 0, 1, 1, 1]), np.array([1, 1, 1, 1, 0, 1, 1, 1]))
    sv = (np.array([0, 0, 0, 0, 1, 1, 1, 1]), np.array([0, 0, 0, 0, 1, 1, 1, 1]))
    files = write_workgroup_analysis(str(tmp_path), "scene", "_b4_l4", 0, 4, 2, 2, primary, secondary, pv, sv)
    d = tmp_path / "WorkGroupSize_2_Version_1"
    assert [str(d / f"scene_b4_l4_c0_{k}.txt") for k in
            ("PrimaryWorkGroupWiskerPlot", "SecondaryWorkGroupWiskerPlot", "WorkGroupData",
             "WorkGroupUniqueWork")] == files
    lines = [open(f).read().splitlines() for f in files]

    # tile 0 depths [2,2,3,3]: median depth[2] = 3, avg 2.5, var 0.25, sd 0.5;
    # tile 1 depths [0,2,2,2]: median 2, avg 1.5, var (2.25+3*0.25)/4 = 0.75,
    # sd 0.866025; sorted by median: tile 1 first
    assert lines[0] == ["median, min, max, lowerStdDeviation, upperStdDeviation",
                        "2, 0, 2, 1.5, 0.866025", "3, 2, 3, 2.5, 0.5"]
    # secondary: tile 0 depths [0,0,0,0], tile 1 [2,2,2,2]
    assert lines[1][1:] == ["0, 0, 0, 0, 0", "2, 2, 2, 2, 0"]

    assert lines[2][0].startswith("stepId, avgPrimaryNodeWork, avgPrimaryNodeUnique,")
    assert lines[2][1:] == [
        # j=0: node work (4+3)/2, unique (1+1)/2, terminations (0+1)/2;
        # secondary node work (0+4)/2, terminations (4+0)/2
        "0, 3.5, 1, 0, 0, 0.5, 1, 1, 0, 0, 2, 1, 0, 0, 2, 1, 1, 0, 0",
        # j=1: secondary only tile 1 is still working: 4/1; terminations stay
        # cumulative (0.5, 2)
        "1, 3.5, 1, 0, 0, 0.5, 1, 1, 0, 0, 4, 1, 0, 0, 2, 1, 1, 0, 0",
        # j=2: node work (2+0)/2 = 1, unique 2 over the one nonzero tile,
        # leaf work 5/2, leaf unique (1+2)/2, terminations 0.5 + 5/2 = 3
        "2, 1, 2, 2.5, 1.5, 3, 2, 2, 2, 1, 0, 0, 4, 3, 4, 0, 0, 3, 3",
        # j=3: only primary tile 0 is left; secondary averages fall to 0
        "3, 0, 0, 2, 2, 4, 0, 0, 2, 2, 0, 0, 0, 0, 4, 0, 0, 0, 0",
    ]

    assert lines[3][0].startswith("loadedPrimaryNodes, loadedPrimaryLeafs,")
    assert lines[3][1:] == [
        # primary nodes 1+1+2+2, leaves 4, max 2/1, min 1/1; secondary all 0
        # (min stays at 0); per-step unique sums: primary 1+1+2+0, 0+0+1+2,
        # secondary 1, 0
        "6, 4, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 4, 3, 1, 0",
        "3, 3, 1, 1, 0, 0, 4, 4, 1, 1, 1, 1, 2, 2, 2, 3",
    ]


def test_cpp_writers_match_python(tmp_path, rt):
    """rt::write_info and rt::write_workgroup_analysis (the C++ host layer)
    write the same bytes as the Python writers from the same counters: a
    3-tile image with two ambient samples and random step records, and
    random histogram counters."""
    import os
    import subprocess

    from bvhraytracer_amd.raytracer import write_info, write_workgroup_analysis

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rng = np.random.default_rng(7)
    W, H, wg = 12, 4, 4  # 3 tiles of 16 rays
    wgsq, tiles = wg * wg, 3

    def tile_records(started, visits):
        steps = int(visits.max()) if started.any() else 0
        rec = np.zeros(steps + 1, STEP)
        rec[0] = (int(started.sum()), 1, 0, 0, int((~started).sum()))
        for t in range(1, steps + 1):
            act = visits >= t
            leaf = rng.random(act.sum()) < 0.4
            nw, lw = int((~leaf).sum()), int(leaf.sum())
            rec[t] = (nw, rng.integers(1, nw + 1) if nw else 0, lw, rng.integers(1, lw + 1) if lw else 0,
                      int((visits == t).sum()))
        return rec

    prim, sec = [], []
    for _ in range(tiles):
        prim.append(tile_records(np.ones(wgsq, bool), rng.integers(1, 30, wgsq)))
        sec.append(np.concatenate([tile_records(rng.random(wgsq) < 0.7, rng.integers(1, 12, wgsq))
                                   for _ in range(2)]))
    pv = (rng.integers(0, 40, tiles * wgsq).astype(np.uint32), rng.integers(0, 20, tiles * wgsq).astype(np.uint32))
    sv = (rng.integers(0, 1500, tiles * wgsq).astype(np.uint32), rng.integers(0, 9, tiles * wgsq).astype(np.uint32))
    py_files = write_workgroup_analysis(str(tmp_path / "py"), "s", "_b8_l4", 1, W, H, wg, prim, sec, pv, sv)

    from bvhraytracer_amd._lib import RtTraceHist

    def hist():
        h = RtTraceHist()
        for f, ctype in RtTraceHist._fields_:
            v = getattr(h, f)
            if hasattr(v, "_length_"):
                for i in range(len(v)):
                    v[i] = int(rng.integers(0, 5000))
            else:
                setattr(h, f, int(rng.integers(20000, 90000)))
        h.rays = 3000
        return h

    hp, hs = hist(), hist()
    as_dict = lambda h: {f: (np.array(getattr(h, f)[:], np.int64) if hasattr(getattr(h, f), "_length_")
                             else int(getattr(h, f))) for f, _ in RtTraceHist._fields_}
    write_info(str(tmp_path / "py_Info.txt"), "s", 8, 4, as_dict(hp), as_dict(hs), bvh_depth=11)

    # the same inputs for the C++ writers, as raw files
    blob = tmp_path / "in.bin"
    with open(blob, "wb") as f:
        f.write(bytes(hp) + bytes(hs))
        for t in prim + sec:
            f.write(np.uint32(len(t)).tobytes() + t.tobytes())
        for a in pv + sv:
            f.write(a.tobytes())
    src = tmp_path / "w.cpp"
    src.write_text(r'''
#include <cstdio>
#include <fstream>
#include "rt_renderer.hpp"
int main(int argc, char** argv) {
  std::ifstream f(argv[1], std::ios::binary);
  rt_trace_hist hp, hs;
  f.read((char*)&hp, sizeof hp); f.read((char*)&hs, sizeof hs);
  std::vector<std::vector<rt_step_stat>> tiles[2];
  for (int k = 0; k < 2; ++k)
    for (int t = 0; t < 3; ++t) {
      uint32_t n; f.read((char*)&n, 4);
      std::vector<rt_step_stat> r(n); f.read((char*)r.data(), n * sizeof(rt_step_stat));
      tiles[k].push_back(r);
    }
  rt::TileVisits pv, sv;
  for (auto* v : {&pv.nodes, &pv.leaves, &sv.nodes, &sv.leaves}) { v->resize(48); f.read((char*)v->data(), 48 * 4); }
  rt::write_workgroup_analysis(std::string(argv[2]) + "/cpp", "s", "_b8_l4", 1, 12, 4, 4, tiles[0], tiles[1], pv, sv);
  rt::write_info(std::string(argv[2]) + "/cpp_Info.txt", "s", 8, 4, 11, hp, hs);
  return 0;
}
''')
    exe = tmp_path / "w"
    lib = os.path.join(root, "bvhraytracer_amd", "librt.so")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(root, "include"), str(src), lib,
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", str(exe)], check=True)
    subprocess.run([str(exe), str(blob), str(tmp_path)], check=True)
    for pf in py_files:
        cf = pf.replace(os.sep + "py" + os.sep, os.sep + "cpp" + os.sep)
        text = open(pf).read()
        assert len(text.splitlines()) > 3 and text == open(cf).read(), os.path.basename(pf)
    assert (tmp_path / "py_Info.txt").read_text() == (tmp_path / "cpp_Info.txt").read_text()
