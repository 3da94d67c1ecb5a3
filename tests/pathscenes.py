"""Scenes whose rendered mean the mathematics fixes, in the reference's .prscene.json format.

They pin the path loop (reference src/integrators/path.cpp:113-238) as a whole, not its parts:
- white furnaces: every surface and medium conserves energy under a uniform white environment of
  unit luminance, so every pixel's expected XYZ is the environment's, the D65 white point at Y = 1.
  A dropped emission-MIS term, a wrong eta-scale, a biased spectral-MIS ratio average or a wrong
  Russian-roulette weight moves the mean off 1. The diffuse furnace is the reference's own test
  (tests/test_render.cpp:50-57, 85-88: mean Y = 1 within 1%); the glass, dispersive-glass,
  chromatic-medium and chromatic-grid furnaces extend it to the refraction, dispersion and
  spectral-tracking branches of the same loop.
- an emissive enclosure: a camera inside a box of reflectance rho that emits Le sees
  L = Le / (1 - rho) everywhere (the Neumann series of the rendering equation), which pins
  area-light next-event estimation and emission MIS with the light sampler's PMF.
- the fog pair of tests/test_media.cpp:59-71, 96-115: a grid medium of constant density renders
  like the equivalent homogeneous medium (within 3%).
"""
import json
import os
import struct

import numpy as np

# XYZ of the D65 white point at Y = 1 (CIE 15:2004, 2-degree observer).
D65_XYZ = (0.95047, 1.0, 1.08883)

_FILM = {"width": 24, "height": 24, "filter": "box"}
_CAM = {"cam": {"type": "pinhole", "position": [0, 0, 4], "look_at": [0, 0, 0], "fov_y": 40}}
_SKY = {"sky": {"type": "environment", "color": [1, 1, 1], "intensity": 1}}


def _furnace(material, geometry=None, extra=None):
    doc = {
        "materials": {"m": material},
        "geometry": {"ball": geometry or {"type": "sphere", "radius": 1}},
        "objects": [{"geometry": "ball", "material": "m"}],
        "lights": dict(_SKY),
        "cameras": dict(_CAM),
        "render": {"camera": "cam", "film": dict(_FILM)},
    }
    if extra:
        doc.update(extra)
    return doc


def write_vol(path, n, value, bmin, bmax):
    """A .vol density grid (format of the reference's VoxelGridField::LoadVol, src/media/medium.cpp:130-166):
    'VOL', version 3, encoding 1 (float32), resolution, 1 channel, bounds, then n^3 values."""
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]))
        f.write(struct.pack("<5i", 1, n, n, n, 1))
        f.write(struct.pack("<6f", bmin, bmin, bmin, bmax, bmax, bmax))
        f.write(np.full(n * n * n, value, np.float32).tobytes())


def furnace_docs(tmpdir):
    """name -> scene document of every furnace scene (each has expected mean XYZ = D65_XYZ)."""
    vol = os.path.join(tmpdir, "unit_density.vol")
    write_vol(vol, 8, 1.0, -1.0, 1.0)
    chroma = {"sigma_s": [1.5, 1.0, 0.6], "sigma_a": 0.0, "g": 0.6}
    return {
        # the reference's kFurnace
        "diffuse": _furnace({"type": "diffuse", "reflectance": 1.0}),
        # smooth glass: reflection + refraction, radiance scaled by 1/eta^2 in and eta^2 out
        "glass": _furnace({"type": "dielectric", "ior": 1.5}),
        # dispersive glass: secondary wavelengths terminate at the first refraction
        "glass_bk7": _furnace({"type": "dielectric", "ior": "BK7"}),
        # a non-absorbing, chromatic, forward-scattering medium behind an index-matched boundary
        "medium": _furnace({"type": "interface", "medium": "fog"},
                           extra={"media": {"fog": dict(type="homogeneous", **chroma)}}),
        # the same medium as a constant-density grid (spectral tracking against the majorant)
        "grid": _furnace({"type": "interface", "medium": "fog"},
                         extra={"media": {"fog": dict(type="grid", file=vol, **chroma)}}),
    }


def enclosure_doc(rho=0.5, le=0.5):
    """Camera at the centre of a 4 x 4 x 4 box (normals flipped inwards) whose twelve triangles are
    Lambertian with reflectance rho and emit two-sided white light of luminance le: expected
    Y = le / (1 - rho). Every triangle is its own light. (A sphere would not do: the reference's
    Shape::PDF, src/geometry/shape.cpp:26-37, intersects from the unoffset point of the previous
    vertex, so a BSDF ray leaving a sphere's inside meets its own origin first and emission MIS
    weights the far side's light with a PDF of about 0; the oracle keeps that behaviour.)"""
    return {
        "materials": {"wall": {"type": "diffuse", "reflectance": rho,
                               "emission": {"color": [1, 1, 1], "luminance": le, "two_sided": True}}},
        "geometry": {"room": {"type": "box", "size": [4, 4, 4]}},
        "objects": [{"geometry": "room", "material": "wall", "flip_normals": True}],
        "cameras": {"cam": {"type": "pinhole", "position": [0.3, -0.2, 0.1], "look_at": [0, 0, -1], "fov_y": 60}},
        "render": {"camera": "cam", "film": dict(_FILM)},
    }


def fog_docs(tmpdir):
    """The reference's FogScene (tests/test_media.cpp:59-71) with the homogeneous medium and with the
    equivalent constant-density grid (8^3 cells of density 1 over [-1.5, 1.5]^3, translated to the ball)."""
    vol = os.path.join(tmpdir, "const2.vol")
    write_vol(vol, 8, 1.0, -1.5, 1.5)

    def fog(medium):
        return {
            "media": {"fog": medium},
            "materials": {"boundary": {"type": "interface", "medium": "fog"},
                          "floor": {"type": "diffuse", "reflectance": 0.6}},
            "geometry": {"ball": {"type": "sphere", "radius": 1}, "ground": {"type": "rect", "size": [8, 8]}},
            "objects": [{"geometry": "ball", "material": "boundary", "transform": {"translate": [0, 1, 0]}},
                        {"geometry": "ground", "material": "floor"}],
            "lights": {"key": {"type": "rect", "position": [0, 4, 1], "look_at": [0, 0, 0], "size": [1, 1],
                               "luminance": 10},
                       "sky": {"type": "environment", "color": [0.3, 0.4, 0.6]}},
            "cameras": {"cam": {"type": "pinhole", "position": [0, 1.2, 4.5], "look_at": [0, 1, 0], "fov_y": 40}},
            "render": {"camera": "cam", "film": dict(_FILM)},
        }

    base = {"sigma_s": [1.5, 1.0, 0.6], "sigma_a": 0.2, "g": 0.3}
    return {"homogeneous": fog(dict(type="homogeneous", **base)),
            "grid": fog(dict(type="grid", file=vol, transform={"translate": [0, 1, 0]}, **base))}


def load(host, tmpdir, name, doc):
    path = os.path.join(tmpdir, name + ".prscene.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return host.load_scene(path)


def mean_and_se(film_sum, film_sq, spp):
    """Image mean per XYZ channel and its standard error, from per-pass FP64 sums (box filter, so
    pixels are independent)."""
    m = film_sum / spp
    var = np.maximum(film_sq / spp - m * m, 0) / spp
    n = m.shape[0] * m.shape[1]
    return m.reshape(-1, 3).mean(0), np.sqrt(var.reshape(-1, 3).sum(0)) / n
