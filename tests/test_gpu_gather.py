"""Load-pipeline calibration (rt_measure_gather): the rate is positive, a
wave with fewer active lanes issues its loads no slower than a full one, and
bad arguments are refused."""
import pytest

from bvhraytracer_amd.host import measure_gather

pytestmark = pytest.mark.gpu


def test_gather_rates(rt):
    full = measure_gather(64)
    part = measure_gather(26)
    l2 = measure_gather(64, 4 << 20)
    print(f"wave loads/s: 64 lanes {full:.3e}, 26 lanes {part:.3e}, 64 lanes over 4 MB {l2:.3e}", flush=True)
    assert full > 1e9 and part > 1e9 and l2 > 1e8
    assert part >= 0.9 * full


@pytest.mark.parametrize("lanes,size", [(0, 16384), (65, 16384), (64, 4096), (64, 3 << 13), (64, 1 << 31)])
def test_gather_arguments(rt, lanes, size):
    with pytest.raises(rt.RtError):
        measure_gather(lanes, size)
