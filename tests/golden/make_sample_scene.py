"""Writes tests/golden/sample_scene.npz: the reference's sample scene (scenes/sphere-pyramid, 14
spheres of every material family on a checker floor) as this build's loader flattens it, with the
studio rig and with the studio + fog rigs, at the scene file's own 960 x 540 film. Used by
`bench.py --sample-scene` to time the scene the reference publishes GPU ms/spp for. Run in the
build container, where the reference tree is present:

    python3 tests/golden/make_sample_scene.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "ar-renderer_amd"))
import prender_amd as pa  # noqa: E402

SCENE = "/root/reference/scenes/sphere-pyramid/scene.prscene.json"


def main():
    host = pa.Host()
    out = {}
    for name, rigs in (("studio", ["studio"]), ("studio_fog", ["studio", "fog"])):
        s = host.load_scene(SCENE, rigs=rigs)
        for f in ("prims", "materials", "spectra", "knots", "lights", "media", "grid"):
            out[f"{name}_{f}"] = getattr(s, f)
        out[f"{name}_camera"] = np.frombuffer(bytes(s.camera), np.uint8)
        out[f"{name}_camera_medium"] = np.int32(s.desc.cameraMedium)
        out[f"{name}_max_depth"] = np.int32(s.defaults["maxDepth"])
        out[f"{name}_seed"] = np.uint64(s.defaults["seed"])
        print(name, len(s.prims), "primitives,", len(s.lights), "lights, warnings:", s.warnings)
    path = os.path.join(HERE, "sample_scene.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
