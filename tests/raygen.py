"""Ray batches for the closest-hit parity tests."""
import numpy as np

import prender_amd as pa


def camera_rays(oracle_scene, scene, n, seed):
    """Primary rays through random film positions, made by the oracle's camera."""
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, pa.RAY_DTYPE)
    fx = rng.uniform(0, scene.width, n).astype(np.float32)
    fy = rng.uniform(0, scene.height, n).astype(np.float32)
    for i in range(n):
        r = oracle_scene.camera_ray(fx[i], fy[i])
        rays["o"][i] = r[:3]
        rays["d"][i] = r[3:]
    rays["tMax"] = np.inf
    return rays


def random_rays(lo, hi, n, seed, t_max=np.inf):
    """Rays with origins uniform in a box (scene bounds, slightly enlarged) and isotropic directions."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    ext = hi - lo
    o = rng.uniform(lo - 0.05 * ext, hi + 0.05 * ext, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, pa.RAY_DTYPE)
    rays["o"] = o
    rays["d"] = d
    rays["tMax"] = t_max
    return rays


def scene_bounds(scene):
    p = np.concatenate([scene.prims["p0"], scene.prims["p1"], scene.prims["p2"]])
    tri = scene.prims["type"] == pa.PRIM_TRIANGLE
    p = np.concatenate([scene.prims["p0"][tri], scene.prims["p1"][tri], scene.prims["p2"][tri]])
    return p.min(axis=0), p.max(axis=0)
