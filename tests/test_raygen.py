"""Host ray generation and hit shading (rt_camera_rays, rt_hit_surface,
rt_spawn_shadow, rt_spawn_diffuse) against independent float64 models: the
camera tile mapping and basis, the plane re-solve of the hit position, the
shadow-ray construction, and the PCG32 stream + cosine-weighted direction
of the diffuse rays (which also pins the shared sin/cos of a fraction of a
turn).  The GPU versions are checked bit-for-bit against these host
functions in test_gpu_shade.py."""
import math

import numpy as np

M64 = (1 << 64) - 1


class Pcg32:
    """PCG-XSH-RR 64/32 (O'Neill), the generator rt_spawn_diffuse uses."""

    def __init__(self, seed, stream):
        self.state = 0
        self.inc = ((stream << 1) | 1) & M64
        self.next()
        self.state = (self.state + seed) & M64
        self.next()

    def next(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & M64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        return (self.next() >> 8) * 2.0 ** -24


def test_pcg32_known_stream():
    # O'Neill's pcg32 demo: seed 42, stream 54 starts 0xa15c02b7 0x7b47f409 0xba1d3330
    g = Pcg32(42, 54)
    assert [g.next() for _ in range(3)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330]


def _dirs(r):
    return np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)


def _origins(r):
    return np.stack([r["ox"], r["oy"], r["oz"]], 1).astype(np.float64)


def test_camera_rays_model(rt):
    pos, tgt, up, focal = (0.0, 5.0, -14.0), (0.0, 3.0, 0.0), (0.0, 1.0, 0.0), 0.866
    W, H, wg = 64, 40, 16  # partial last tile row (40 = 2.5 tiles)
    rays = rt.camera_rays(pos, tgt, W, H, wg)
    assert len(rays) == W * H
    p = np.array(pos, np.float64)
    f = np.array(tgt, np.float64) - p
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    k = 0
    exp = []
    for i in range((W // wg) * ((H + wg - 1) // wg)):
        for j in range(wg * wg):
            row = (i * wg) // W * wg + j // wg
            if row >= H:
                continue
            w = int(((i * wg) % W) - W / 2 + (j % wg))
            h = int(-(((i * wg) // W) * wg - H / 2) - (j // wg))
            # camera space looks down -z; the camera-to-world columns are s, u, -f
            exp.append(s * (w / W) + u * (h / W) + f * focal)
            k += 1
    exp = np.array(exp)
    got = _dirs(rays)
    assert np.abs(got - exp).max() < 1e-5
    assert np.all(_origins(rays) == p)
    assert np.all(rays["tmax"] == np.float32(222222.0))
    # inverse direction: 1/d, FLT_MAX where d == 0
    d = rays["dx"]
    assert np.all(rays["ix"][d != 0] == (np.float32(1) / d[d != 0]))
    # the centre pixels look along the view direction
    assert abs(np.dot(got[0] / np.linalg.norm(got[0]), f)) > 0.5


def test_shadow_rays_model(rt):
    rng = np.random.default_rng(1)
    n = 500
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pos = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
    hits = np.zeros(n, rt.HIT_DTYPE)
    hits["slot"] = np.where(rng.random(n) < 0.8, 0, -1)
    light = (0.0, 9.5, 0.0)
    rays, src = rt.spawn_shadow(nrm, pos, hits, light, 1.0)
    assert np.array_equal(src, np.nonzero(hits["slot"] >= 0)[0])
    o = pos[src].astype(np.float64) + nrm[src].astype(np.float64) * 0.001
    assert np.abs(_origins(rays) - o).max() < 1e-5
    assert np.abs(_dirs(rays) - (np.array(light) - o)).max() < 1e-4
    assert np.all(rays["tmax"] == np.float32(1.0))


def test_diffuse_rays_model(rt):
    rng = np.random.default_rng(2)
    n = 400
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pos = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
    prim = rt.make_rays(np.zeros((n, 3), np.float32), rng.normal(size=(n, 3)).astype(np.float32))
    hits = np.zeros(n, rt.HIT_DTYPE)
    hits["slot"] = np.where(rng.random(n) < 0.7, 1, -1)
    hit_idx = np.nonzero(hits["slot"] >= 0)[0]
    count, seed = 3000, 7
    rays = rt.spawn_diffuse(prim, nrm, pos, hits, count, seed=seed, tmax=100.0)
    got_d, got_o = _dirs(rays), _origins(rays)
    pd = _dirs(prim)
    for r in range(count):
        i = hit_idx[r % len(hit_idx)]
        nn = nrm[i].astype(np.float64)
        if np.dot(nn, pd[i]) > 0:
            nn = -nn
        g = Pcg32(seed, r)
        u1, u2 = g.uniform(), g.uniform()
        phi = 2 * math.pi * u2
        lx, ly, lz = math.sqrt(u1) * math.cos(phi), math.sqrt(u1) * math.sin(phi), math.sqrt(1 - u1)
        sgn = math.copysign(1.0, nn[2])
        a = -1.0 / (sgn + nn[2])
        b = nn[0] * nn[1] * a
        t = np.array([1 + sgn * nn[0] ** 2 * a, sgn * b, -sgn * nn[0]])
        bt = np.array([b, sgn + nn[1] ** 2 * a, -nn[1]])
        d = t * lx + bt * ly + nn * lz
        assert np.abs(got_d[r] - d).max() < 2e-5, (r, got_d[r], d)
        assert np.abs(got_o[r] - (pos[i] + nn * 0.001)).max() < 1e-5
    assert np.all(rays["tmax"] == np.float32(100.0))
    # cosine-weighted about the face-forwarded normal: mean cos = 2/3
    nn = nrm[hit_idx[np.arange(count) % len(hit_idx)]].astype(np.float64)
    flip = np.einsum("ij,ij->i", nn, pd[hit_idx[np.arange(count) % len(hit_idx)]]) > 0
    nn[flip] *= -1
    cosv = np.einsum("ij,ij->i", got_d, nn)
    assert cosv.min() >= 0 and abs(cosv.mean() - 2 / 3) < 0.02


def test_hit_surface_model(rt):
    tris = rt.interior_scene(3, 3000, 4)
    bvh = rt.Bvh(tris, 8, 8)
    Lm = bvh.leaf_memory
    nblocks = len(bvh.tri_points) // (9 * Lm)
    rng = np.random.default_rng(3)
    hits = np.zeros(300, rt.HIT_DTYPE)
    hits["leaf"] = rng.integers(0, nblocks, 300) * 9 * Lm
    hits["slot"] = 0  # slot 0 of every leaf holds a real triangle
    T = bvh.tri_points
    # planes at stride lc: Lm, or the loop count before the NaN marker that
    # ends a block whose leaf is not full (getLeafLoopCount)
    end = hits["leaf"] + 9 * Lm
    lc = np.where(np.isnan(T[end - 1]), T[end - 2], Lm).astype(np.int64)
    assert set(np.unique(lc)) <= {4, 8} and (lc < Lm).any()
    p0 = np.stack([T[hits["leaf"] + lc * a] for a in range(3)], 1).astype(np.float64)
    p1 = np.stack([T[hits["leaf"] + lc * (3 + a)] for a in range(3)], 1).astype(np.float64)
    p2 = np.stack([T[hits["leaf"] + lc * (6 + a)] for a in range(3)], 1).astype(np.float64)
    full = np.ones(len(hits), bool)
    cen = (p0 + p1 + p2) / 3
    org = cen + rng.normal(size=cen.shape) * 2.0
    rays = rt.make_rays(org.astype(np.float32), cen.astype(np.float32))
    nrm, pos = bvh.surface(rays, hits)
    n64 = np.cross(p1 - p0, p2 - p0)
    n64 /= np.linalg.norm(n64, axis=1, keepdims=True)
    ok = full & (np.abs(np.einsum("ij,ij->i", n64, _dirs(rays))) > 1e-3 * np.linalg.norm(_dirs(rays), axis=1))
    assert ok.sum() > 100
    assert np.abs(nrm[ok] - n64[ok]).max() < 1e-5
    # the position lies on the triangle's plane, on the ray, at the centroid
    assert np.abs(pos[ok] - cen[ok]).max() < 1e-3
    # a miss gives NaN
    miss = hits.copy()
    miss["slot"] = -1
    nm, pm = bvh.surface(rays[:4], miss[:4])
    assert np.isnan(nm).all() and np.isnan(pm).all()


def _mix(index, sample):
    """The (u, v) hash of the ambient directions: a 64-bit mix with
    SplitMix64's finaliser constants, 24 bits each."""
    h = (index * 0x9E3779B97F4A7C15 + sample * 0xBF58476D1CE4E5B9 + 1) & M64
    h ^= h >> 31
    h = (h * 0x94D049BB133111EB) & M64
    h ^= h >> 29
    return (h >> 40) * 2.0 ** -24, ((h >> 16) & 0xFFFFFF) * 2.0 ** -24


def test_ambient_rays_model(rt):
    """rt_spawn_ambient against a float64 model of the uniform sphere point
    (z = 1 - 2u, r = sqrt(1 - z^2), angle 2 pi v) flipped into the normal's
    hemisphere; origin P + n*0.001, direction used as is (updateDirection),
    tMax = dist; misses get no ray.  Over many samples the directions are
    uniform on the hemisphere: mean cosine 1/2."""
    rng = np.random.default_rng(5)
    n = 3000
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pos = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
    hits = np.zeros(n, rt.HIT_DTYPE)
    hits["slot"] = np.where(rng.random(n) < 0.75, 0, -1)
    for sample in (0, 3):
        rays, src = rt.spawn_ambient(nrm, pos, hits, sample, 2.5)
        assert np.array_equal(src, np.nonzero(hits["slot"] >= 0)[0])
        got = _dirs(rays)
        for k, i in enumerate(src[:400]):
            u, v = _mix(int(i), sample)
            z = 1.0 - 2.0 * u
            r = math.sqrt(max(0.0, 1.0 - z * z))
            d = np.array([r * math.cos(2 * math.pi * v), r * math.sin(2 * math.pi * v), z])
            if np.dot(d, nrm[i].astype(np.float64)) <= 0:
                d = -d
            assert np.abs(got[k] - d).max() < 2e-6, (i, got[k], d)
        o = pos[src].astype(np.float64) + nrm[src].astype(np.float64) * 0.001
        assert np.abs(_origins(rays) - o).max() < 1e-5
        assert np.all(rays["tmax"] == np.float32(2.5))
        assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)
        cosv = np.einsum("ij,ij->i", got, nrm[src].astype(np.float64))
        assert cosv.min() >= 0 and abs(cosv.mean() - 0.5) < 0.02
    # (0, 0) by hand: h = 1 -> 1 * 0x94D049BB133111EB, then h ^= h >> 29
    h = 0x94D049BB133111EB
    h ^= h >> 29
    assert _mix(0, 0) == ((h >> 40) * 2.0 ** -24, ((h >> 16) & 0xFFFFFF) * 2.0 ** -24)
