"""_Info.txt writer (write_info): the line keys and per-ray normalisation of
the reference's instrumented renderer (cameraData.h:241-329), read back the
way the reference's Analysis/DataManager.py reads them (find the key in a
line, take the first number on it)."""
import numpy as np

from bvhraytracer_amd.raytracer import write_info


def first_number(lines, key):
    for line in lines:
        if key in line:
            for tok in line.split():
                try:
                    return float(tok)
                except ValueError:
                    pass
    raise KeyError(key)


def stats(rays, nodes, leaves, at, ah, tt, th):
    nd = np.zeros(64, np.int64)
    nd[0], nd[1] = rays, nodes - rays
    ld = np.zeros(64, np.int64)
    ld[2] = leaves
    cf = np.zeros(17, np.int64)
    cf[4] = nodes
    lf = np.zeros(17, np.int64)
    lf[3] = leaves
    return dict(rays=rays, node_visits=nodes, leaf_visits=leaves, aabb_tests=at, aabb_hits=ah, tri_tests=tt,
                tri_hits=th, node_depth=nd, leaf_depth=ld, child_fullness=cf, leaf_fullness=lf)


def test_info_keys_and_values(tmp_path):
    p = stats(100, 300, 150, 1200, 400, 450, 90)
    s = stats(50, 100, 40, 400, 100, 120, 30)
    path = tmp_path / "x_b4_l4_Info.txt"
    write_info(path, "x", 4, 4, p, s)
    lines = path.read_text().splitlines()
    assert lines[0] == "scenario x with branching factor of 4 and leafsize of 4"
    assert first_number(lines, "primary intersections node:") == 3.0
    assert first_number(lines, "primary aabb intersections:") == 12.0
    # std::ostream float format: 6 significant digits
    assert "primary aabb success ration: 0.333333" in lines
    assert "primary waste factor: 0.708333" in lines  # 1 - (150 + 300 - 100) / (300 * 4)
    assert first_number(lines, "primary intersections leaf:") == 1.5
    assert first_number(lines, "primary primitive intersections:") == 4.5
    assert "primary primitive success ratio: 0.2" in lines
    assert first_number(lines, "secondary intersections node:") == 2.0
    assert first_number(lines, "secondary intersections leaf:") == 0.8
    assert "average child fullness: 4.000000" in lines  # std::to_string
    assert "average leaf fullness: 3.000000" in lines
    # child fullness 0..N, leaf fullness 0..L (cameraData.h:179,186)
    i = lines.index("intersections with nodes with x children :")
    assert lines[i + 2:i + 7] == ["0 : 0", "1 : 0", "2 : 0", "3 : 0", "4 : 3"] and lines[i + 7] == ""
    i = lines.index("primary node intersections at depth x :")
    assert lines[i + 1] == "0 : 1" and lines[i + 2] == "1 : 2"
    i = lines.index("secondary leaf intersections at depth x :")
    assert lines[i + 3] == "2 : 0.8"
