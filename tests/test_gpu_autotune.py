"""Per-scene autotuner: candidates are measured on the GPU, the winner is
cached by content hash, and the tuned configuration traces exactly like the
oracle (tuning changes speed, never results)."""
import pytest

from bvhraytracer_amd.autotune import Autotuner
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


def test_autotune_small_scene(rt, tmp_path):
    tris = rt.interior_scene(29, 20000, 10)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 64, 16)
    cache = str(tmp_path / "tune.json")
    print("autotune: 2 N/L x 2 blocks x 2 schedules", flush=True)
    at = Autotuner(cache_path=cache, nl=((4, 4), (8, 8)), blocks=(128, 256), windows=(16,), schedules=(0, 16),
                   sample=4096, reps=1)
    best = at.tune_scene(tris, rays)
    assert not best["cached"]
    assert (best["N"], best["L"]) in ((4, 4), (8, 8))
    assert len(best["tried"]) == 2 and best["mrays"] > 0
    assert best["block"] in (128, 256) and best["window"] == 16 and best["schedule"] in (0, 16)
    assert best["exact_boxes"] in (False, True)
    at.save()
    again = Autotuner(cache_path=cache).tune_scene(tris, rays)
    assert again["cached"] and (again["N"], again["L"], again["schedule"]) == (best["N"], best["L"], best["schedule"])

    bvh = rt.Bvh(tris, best["N"], best["L"])
    scene = rt.Scene(bvh)
    scene.set_launch(best["block"], best["window"])
    scene.set_schedule(best["schedule"])
    scene.exact_boxes = best["exact_boxes"]
    g, gc = scene.intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "tuned")
