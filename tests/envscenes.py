"""Image-environment scenes for the environment-map tests: PFM / PPM writers and scene documents in
the reference's .prscene.json format (lights of type "environment" with a "file", and "sun_sky")."""
import json
import os

import numpy as np


def write_pfm(path, rgb):
    """rgb (h, w, 3) float, top row first -> PFM (little-endian, bottom row first)."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(rgb[::-1], dtype="<f4").tobytes())


def write_ppm(path, rgb8):
    h, w, _ = rgb8.shape
    with open(path, "wb") as f:
        f.write(f"P6\n# test map\n{w} {h}\n255\n".encode())
        f.write(np.ascontiguousarray(rgb8, dtype=np.uint8).tobytes())


def studio_map(w=64, h=32, seed=3):
    """A sky-like map: a blue-to-white vertical gradient, a warm ground, a bright small 'sun' blob and
    some coloured noise, so that sampling, bilinear lookups and the RGB uplift all matter."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    img = np.zeros((h, w, 3), np.float32)
    sky = np.stack([0.3 + 0.5 * v, 0.45 + 0.4 * v, 0.9 - 0.2 * v], 1)
    img[:] = sky[:, None, :]
    img[h // 2:] = np.array([0.35, 0.25, 0.15])
    uu, vv = np.meshgrid(u, v)
    blob = np.exp(-(((uu - 0.3) / 0.03) ** 2 + ((vv - 0.25) / 0.05) ** 2))
    img += (40.0 * blob)[..., None] * np.array([1.0, 0.9, 0.7])
    img *= 1 + 0.2 * rng.random((h, w, 1))
    return img.astype(np.float32)


def constant_map(rgb, w=16, h=8):
    return np.broadcast_to(np.asarray(rgb, np.float32), (h, w, 3)).copy()


def env_doc(map_path, objects=True, rotation=0.0, intensity=1.0, film=(24, 24)):
    doc = {
        "lights": {"env": {"type": "environment", "file": map_path, "intensity": intensity, "rotation": rotation}},
        "cameras": {"cam": {"type": "pinhole", "position": [0, 0.5, 4], "look_at": [0, 0, 0], "fov_y": 50}},
        "render": {"camera": "cam", "film": {"width": film[0], "height": film[1], "filter": "box"}},
    }
    if objects:
        doc["materials"] = {"white": {"type": "diffuse", "reflectance": 0.8},
                            "gold": {"type": "conductor", "metal": "Au", "roughness": 0.3},
                            "glass": {"type": "dielectric", "ior": 1.5}}
        doc["geometry"] = {"ball": {"type": "sphere", "radius": 0.6}, "floor": {"type": "rect", "size": [6, 6]}}
        doc["objects"] = [{"geometry": "floor", "material": "white", "transform": {"translate": [0, -0.6, 0]}},
                          {"geometry": "ball", "material": "gold", "transform": {"translate": [-0.7, 0, 0]}},
                          {"geometry": "ball", "material": "glass", "transform": {"translate": [0.7, 0, 0]}}]
    return doc


def furnace_env_doc(map_path):
    """A reflectance-1 diffuse sphere under an image environment."""
    return {
        "materials": {"m": {"type": "diffuse", "reflectance": 1.0}},
        "geometry": {"ball": {"type": "sphere", "radius": 1}},
        "objects": [{"geometry": "ball", "material": "m"}],
        "lights": {"env": {"type": "environment", "file": map_path, "intensity": 1.0}},
        "cameras": {"cam": {"type": "pinhole", "position": [0, 0, 4], "look_at": [0, 0, 0], "fov_y": 40}},
        "render": {"camera": "cam", "film": {"width": 24, "height": 24, "filter": "box"}},
    }


def sun_sky_doc(resolution=64):
    return {
        "materials": {"white": {"type": "diffuse", "reflectance": 0.7}},
        "geometry": {"ball": {"type": "sphere", "radius": 0.6}, "floor": {"type": "rect", "size": [6, 6]}},
        "objects": [{"geometry": "floor", "material": "white", "transform": {"translate": [0, -0.6, 0]}},
                    {"geometry": "ball", "material": "white"}],
        "lights": {"sky": {"type": "sun_sky", "elevation": 35, "azimuth": 40, "turbidity": 3,
                           "resolution": resolution}},
        "cameras": {"cam": {"type": "pinhole", "position": [0, 0.8, 4], "look_at": [0, 0, 0], "fov_y": 50}},
        "render": {"camera": "cam", "film": {"width": 24, "height": 24, "filter": "box"}},
    }


def load(host, tmpdir, name, doc):
    path = os.path.join(tmpdir, name + ".prscene.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return host.load_scene(path)
