"""MI355X-native wide-BVH ray-query engine (drop-in for the traversal path of
Hengoo/BVHRaytracer).  The compute path is librt.so (HIP for gfx950); this
package is its Python host binding."""
from ._lib import (COUNTS_DTYPE, HIT_DTYPE, RAY_DTYPE, STEP_DTYPE, RtError, lib)  # noqa: F401
from .host import (Batch, Bvh, FastNodeArrays, Scene, Server, camera_rays, camera_rays_device,  # noqa: F401
                   generate_scene, interior_scene, load_sah_factor, make_rays, spawn_ambient, spawn_diffuse,
                   spawn_diffuse_device,
                   spawn_shadow, spawn_shadow_device, spheres_scene)

__all__ = ["Batch", "Bvh", "Server", "FastNodeArrays", "Scene", "camera_rays", "camera_rays_device", "generate_scene",
           "interior_scene", "load_sah_factor", "make_rays", "spawn_ambient", "spawn_diffuse", "spawn_diffuse_device", "spawn_shadow",
           "spawn_shadow_device", "spheres_scene", "RAY_DTYPE",
           "HIT_DTYPE", "COUNTS_DTYPE", "STEP_DTYPE", "RtError", "lib"]

lib()  # load librt.so now: a missing build fails at import, never later
