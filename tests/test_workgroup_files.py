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
