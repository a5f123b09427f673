"""The leaf test's reciprocal (inv = 1.f / div, rayTracerTemplate.ispc:59).
The kernels form it from the hardware reciprocal refined by one Newton step
and rely on it being the correctly rounded quotient for every divisor.
rt_check_recip runs that exact device function:
  * over all 2^32 float bit patterns against the device's correctly rounded
    division (one bounded kernel, mismatches counted on the device);
  * on sampled ranges against the host's IEEE float32 division (numpy), which
    pins the device division itself: around 1, both ends of the Newton-step
    range [2^-126, 2^126), denormals, signed zeros, huge values, inf, NaN."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(rt, first, count, values=None):
    lib = rt.lib()
    mm, lo = ctypes.c_uint64(), ctypes.c_uint64()
    rt._lib.check(lib.rt_check_recip(first, count, None if values is None else values.data_ptr(),
                                     ctypes.byref(mm), ctypes.byref(lo), None))
    return mm.value, lo.value


def test_recip_equals_division_for_every_float(rt):
    print("rt_check_recip over all 2^32 patterns", flush=True)
    mm, lo = _check(rt, 0, 1 << 32)
    assert mm == 0, f"{mm} mismatches, lowest pattern {lo:#010x}"
    assert lo == (1 << 64) - 1


@pytest.mark.parametrize("first", [
    0x3F000000,  # [0.5, 2): around 1
    0x00800000,  # 2^-126: lowest exponent of the Newton-step range
    0x7E800000,  # 2^126: first exponent outside it (full division)
    0x7E000000,  # 2^125 ... just inside
    0x00000000,  # +0 and the positive denormals
    0x80000000,  # -0 and the negative denormals
    0xBF000000,  # negative, around -1
    0x7F780000,  # up to +inf and into the NaNs
    0xFF780000,  # -inf and negative NaNs
])
def test_recip_matches_host_ieee_division(rt, first):
    import torch

    count = 1 << 20
    print(f"rt_check_recip range {first:#010x} + {count}", flush=True)
    vals = torch.empty(count, dtype=torch.float32, device="cuda")
    mm, _ = _check(rt, first, count, vals)
    assert mm == 0
    bits = (np.uint64(first) + np.arange(count, dtype=np.uint64)) & np.uint64(0xFFFFFFFF)
    x = bits.astype(np.uint32).view(np.float32)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        want = np.float32(1.0) / x
    got = vals.cpu().numpy()
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, f"{bad.size} differ from IEEE division, first x = {x[bad[0]]!r} ({bits[bad[0]]:#x})"
