"""The _Perf.png path of CameraFast (cameraFast.h:404-439 gray values,
:829-852 tile-order reordering and PNG): the PNG is checked with an
independent decoder (chunk CRCs, IHDR, rows), the reordering against the
geometry of camera_rays, and the gray values against hand-derived float32
results."""
import struct
import zlib

import numpy as np
import pytest


def decode_png(data):
    """Minimal PNG reader for what write_png emits: 8-bit RGBA, no interlace,
    every row filter type 0.  Verifies the signature and every chunk CRC."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF, tag
        chunks.append((tag, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1] == (b"IEND", b"")
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    assert len(raw) == h * (1 + 4 * w)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip(tmp_path, rt):
    from bvhraytracer_amd.raytracer import write_png

    rng = np.random.default_rng(0)
    for h, w in ((1, 1), (7, 13), (64, 48)):
        img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
        p = tmp_path / f"i{h}x{w}.png"
        write_png(str(p), img)
        assert np.array_equal(decode_png(p.read_bytes()), img)
    with pytest.raises(ValueError):
        write_png(str(tmp_path / "bad.png"), np.zeros((4, 4, 3), np.uint8))


@pytest.mark.parametrize("W,H,wg", [(64, 48, 16), (64, 40, 16), (48, 36, 8), (32, 6, 16)])
def test_tile_order_to_rows_geometry(rt, W, H, wg):
    """After reordering, pixel (h, w) holds the ray of that pixel: its offset
    along the camera's right vector grows with w and along its up vector
    falls with h, by one pixel (1/W) per step."""
    from bvhraytracer_amd.raytracer import tile_order_to_rows

    pos, tgt = (0.0, 5.0, -14.0), (0.0, 3.0, 0.0)
    rays = rt.camera_rays(pos, tgt, W, H, wg)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    img = tile_order_to_rows(d, W, H, wg)
    assert img.shape == (H, W, 3)
    f = np.array(tgt) - np.array(pos)
    f /= np.linalg.norm(f)
    s = np.cross(f, (0.0, 1.0, 0.0))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    x, y = img @ s, img @ u
    assert np.allclose(np.diff(x, axis=1), 1.0 / W, atol=1e-5)
    assert np.allclose(np.diff(y, axis=0), -1.0 / W, atol=1e-5)
    assert np.allclose(x[:, W // 2], 0.0, atol=1e-5)
    with pytest.raises(ValueError):
        tile_order_to_rows(d[:-1], W, H, wg)


def test_ambient_gray_values(rt):
    """(uint8)((1 - k / n) * 255) in float32.  For n = 3: 1/3 rounds up to
    11184811 * 2^-25, 1 - 1/3 is a tie between float32 neighbours and rounds
    to even (11184810 * 2^-24), times 255 rounds below 170 -> 169; likewise
    1 - 2/3 = 0.33333331 times 255 rounds below 85 -> 84 (float64 would give
    170 and 85).  Misses are 0; with no samples every pixel is 0."""
    from bvhraytracer_amd.raytracer import ambient_image

    occ = np.array([0, 1, 2, 3, 0, 3], np.uint16)
    hit = np.array([1, 1, 1, 1, 0, 0], bool)
    assert ambient_image(occ, hit, 3).tolist() == [255, 169, 84, 0, 0, 0]
    assert ambient_image(np.array([0, 1, 2, 3, 4, 5], np.uint16), np.ones(6, bool), 5).tolist() == \
        [255, 204, 153, 101, 50, 0]
    assert ambient_image(occ, hit, 0).tolist() == [0] * 6
