"""The synthetic workloads of the benchmark configurations, shared by the
benchmark and the tests so both trace the same scenes and rays.

* headline: the 1M-triangle clustered interior (seed 2), 1920x1080 primary
  rays from (0, 5, -14) to (0, 3, 0), their shadow rays to the light at
  (0, 9.5, 0), and cosine-distributed diffuse bounces from their hits;
* batch: 64 interior scenes (seeds 100..163) with lognormal triangle and ray
  counts, half primary rays of a random camera, half diffuse bounces;
* serving: 16 resident interior scenes (100k..2M triangles, log-spaced, seeds
  200..215) and Poisson request arrivals.

The reference renders glTF scenes (modelLoader.h); these procedural scenes
stand in for them at the sizes the benchmark is quoted on.
"""
import numpy as np

from . import host as rt

CAMERA_POS = (0.0, 5.0, -14.0)
CAMERA_TARGET = (0.0, 3.0, 0.0)
LIGHT = (0.0, 9.5, 0.0)


def clamp_lognormal(rng, mu, sigma, lo, hi, size):
    return np.clip(np.exp(rng.normal(mu, sigma, size)), lo, hi).astype(np.int64)


def random_camera_rays(rng, n):
    """Primary rays of a random camera inside the 30x10x30 room: n rays of a
    256-wide image (tile order), the last tile row cut to n."""
    pos = rng.uniform([-12.0, 1.0, -12.0], [12.0, 9.0, 12.0]).astype(np.float32)
    d = rng.normal(size=3)
    d[1] *= 0.3
    tgt = (pos + 5.0 * d / np.linalg.norm(d)).astype(np.float32)
    h = max(16, int(np.ceil(n / 256 / 16)) * 16)
    return rt.camera_rays(tuple(pos), tuple(tgt), 256, h, 16)[:n]


def headline(branching=8, leaf=8, diffuse=8_000_000, triangles=1_000_000):
    """(tris, bvh, scene, primary rays, primary hits, shadow rays, diffuse rays)
    of the 1M-triangle interior at N/L; the scene is resident on the GPU."""
    tris = rt.interior_scene(2, triangles)
    bvh = rt.Bvh(tris, branching, leaf)
    scene = rt.Scene(bvh)
    prim = rt.camera_rays(CAMERA_POS, CAMERA_TARGET, 1920, 1080, 16)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    shadow, _ = rt.spawn_shadow(nrm, pos, hits, LIGHT, 1.0)
    rays = rt.spawn_diffuse(prim, nrm, pos, hits, diffuse, seed=3) if diffuse else None
    return tris, bvh, scene, prim, hits, shadow, rays


# ---- batch: 64 scenes in one persistent launch ------------------------------

BATCH_SCENES = 64


def batch_sizes():
    """Triangles clamp(LogNormal(ln 2e5, 1), 1e4, 2e6) and rays
    clamp(LogNormal(ln 6.4e4, 1.2), 4096, 2^20) for the 64 scenes (seed 4)."""
    rng = np.random.default_rng(4)
    ntris = clamp_lognormal(rng, np.log(2e5), 1.0, 1e4, 2e6, BATCH_SCENES)
    nrays = clamp_lognormal(rng, np.log(6.4e4), 1.2, 4096, 1 << 20, BATCH_SCENES)
    return ntris, nrays


def batch_scene(i, branching=8, leaf=8, ntris=None, nrays=None):
    """Scene i of the batch (seed 100 + i): (tris, bvh, scene, rays) with half
    primary rays of a random camera in the room and half diffuse bounces
    from their hits (closest hit).  The scene is resident on the GPU."""
    if ntris is None or nrays is None:
        nt, nr = batch_sizes()
        ntris, nrays = nt[i], nr[i]
    tris = rt.interior_scene(100 + i, int(ntris))
    bvh = rt.Bvh(tris, branching, leaf)
    sc = rt.Scene(bvh)
    npri = int(nrays) // 2
    prim = random_camera_rays(np.random.default_rng(1000 + i), npri)
    ph = sc.intersect(prim)
    nrm, pos = bvh.surface(prim, ph)
    if (ph["slot"] >= 0).any():
        diff = rt.spawn_diffuse(prim, nrm, pos, ph, int(nrays) - npri, seed=100 + i)
        rays = np.concatenate([prim, diff])
    else:
        rays = prim
    return tris, bvh, sc, rays


# ---- serving: 16 resident scenes, Poisson arrivals --------------------------

SERVE_SCENES = 16
SERVE_POOL = 1 << 19  # rays per scene each request draws a window from


def serving_sizes():
    return np.round(np.exp(np.linspace(np.log(1e5), np.log(2e6), SERVE_SCENES))).astype(np.int64)


def serving_scene(i, branching=8, leaf=8):
    """Resident scene i (seed 200 + i): (tris, bvh, scene, closest rays,
    shadow rays).  Closest-hit requests draw windows of 2^20 rays (2^19
    random-camera primaries, then 2^19 diffuse bounces from their hits);
    shadow requests draw windows of 2^19 shadow rays of those hits."""
    tris = rt.interior_scene(200 + i, int(serving_sizes()[i]))
    bvh = rt.Bvh(tris, branching, leaf)
    sc = rt.Scene(bvh)
    prim = random_camera_rays(np.random.default_rng(2000 + i), SERVE_POOL)
    ph = sc.intersect(prim)
    nrm, pos = bvh.surface(prim, ph)
    diff = rt.spawn_diffuse(prim, nrm, pos, ph, SERVE_POOL, seed=200 + i)
    shadow, _ = rt.spawn_shadow(nrm, pos, ph, LIGHT, 1.0)
    shadow = np.resize(shadow, SERVE_POOL)
    return tris, bvh, sc, np.concatenate([prim, diff]), shadow


def serving_requests(rate, seconds, seed=5):
    """Poisson arrivals at `rate` req/s over `seconds`: per request the gap
    before it (s), its ray count clamp(LogNormal(ln 2.3e4, 1.3), 1024, 2^19),
    its scene, whether it is a shadow (any-hit) request, and the offset of
    its window in the scene's ray pool."""
    rng = np.random.default_rng(seed)
    n = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n)
    sizes = clamp_lognormal(rng, np.log(2.3e4), 1.3, 1024, SERVE_POOL, n)
    scene_of = rng.integers(0, SERVE_SCENES, n)
    shadow = rng.random(n) < 0.5
    offs = np.array([int(rng.integers(0, (2 * SERVE_POOL if not shadow[k] else SERVE_POOL) - sizes[k] + 1))
                     for k in range(n)], np.int64)
    return gaps, sizes, scene_of, shadow, offs
