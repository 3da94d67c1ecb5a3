"""Builds the reference's v4 scene description (src/gpu/gpu_api.h) from a flat scene, as the
reference exporter does (src/integrators/gpu_path.cpp:104-456): spectra tabulated at 95
wavelengths, one v4 light per emitter, the constant environment in envSpectrum, the film's
filter with its importance-sampling CDF. Also gives the extension scene that describes the same
tabulated input, for the oracle."""
import ctypes as C

import numpy as np

import prender_amd as pa

LAMBDAS = 360.0 + 5.0 * np.arange(95)


def tabulate(host, scene):
    out = np.zeros(len(scene.spectra), pa.SPECTRUM95_DTYPE)
    for i, row in enumerate(scene.spectra):
        out["v"][i] = [host.spectrum_value(row, scene.knots, float(l)) for l in LAMBDAS]
    return out


def filter_v4(cam):
    """Filter settings plus its 64-bin CDF with a 1e-3 peak floor (gpu_path.cpp:431-456)."""
    f = pa.FilterV4()
    f.type, f.radius = cam.filterType, cam.filterRadius
    if f.type == pa.FILTER_BOX:
        f.radius = 0.5
    f.sigma = f.radius / 3

    def ev(x):
        x = np.abs(x)
        if f.type == pa.FILTER_GAUSSIAN:
            v = np.maximum(0, np.exp(-x * x / (2 * f.sigma ** 2)) - np.exp(-f.radius ** 2 / (2 * f.sigma ** 2)))
        elif f.type == pa.FILTER_BH:
            t = 0.5 + x / (2 * f.radius)
            v = 0.35875 - 0.48829 * np.cos(2 * np.pi * t) + 0.14128 * np.cos(4 * np.pi * t) - 0.01168 * np.cos(
                6 * np.pi * t)
        else:
            v = np.ones_like(x)
        return np.where(x > f.radius, 0, v)

    n = 4096
    xs = -f.radius + (np.arange(n) + 0.5) * (2 * f.radius / n)
    s1 = ev(xs).sum() * 2 * f.radius / n
    f.norm = 1 / (s1 * s1)
    xb = -f.radius + (np.arange(64) + 0.5) * (2 * f.radius / 64)
    bins = np.abs(ev(xb))
    bins = bins + 1e-3 * bins.max()
    cdf = np.concatenate([[0], np.cumsum(bins / bins.sum())])
    cdf[-1] = 1
    f.cdf[:] = cdf.astype(np.float32)
    return f


class V4Scene:
    """The v4 SceneDesc of `scene` and the arrays it points at."""

    def __init__(self, host, scene, sheen_table=None):
        self.spectra = tabulate(host, scene)
        self.sheen = None if sheen_table is None else np.ascontiguousarray(sheen_table, np.float32)
        self.prims = scene.prims
        self.materials = scene.materials
        lights = [l for l in scene.lights if l["kind"] in (pa.LIGHT_AREA, pa.LIGHT_POINT)]
        env = [l for l in scene.lights if l["kind"] == pa.LIGHT_ENV]
        if len(lights) + len(env) != len(scene.lights) or len(env) > 1:
            raise ValueError("scene has lights the v4 interface cannot express")
        self.lights = np.zeros(len(lights), pa.LIGHT_V4_DTYPE)
        for i, l in enumerate(lights):
            v = self.lights[i]
            v["type"] = pa.V4_LIGHT_AREA if l["kind"] == pa.LIGHT_AREA else pa.V4_LIGHT_POINT
            v["primitive"], v["spectrum"], v["scale"] = l["primitive"], l["spectrum"], l["scale"]
            v["position"], v["twoSided"], v["cosPower"] = l["position"], l["twoSided"], l["cosPower"]
            v["pmf"] = 1.0 / len(lights)
        if np.any(scene.media["kind"] != pa.MEDIUM_HOMOGENEOUS):
            raise ValueError("v4 media are homogeneous only")
        self.media = np.zeros(len(scene.media), pa.MEDIUM_V4_DTYPE)
        for f in ("sigmaA", "sigmaS", "scale", "g"):
            self.media[f] = scene.media[f]
        d = self.desc = pa.SceneDescV4()
        d.apiVersion = 4
        d.prims, d.primCount = scene.prims.ctypes.data, len(scene.prims)
        d.materials, d.materialCount = scene.materials.ctypes.data, len(scene.materials)
        d.lights = self.lights.ctypes.data if len(self.lights) else None
        d.lightCount = len(self.lights)
        d.spectra, d.spectrumCount = self.spectra.ctypes.data, len(self.spectra)
        d.media = self.media.ctypes.data if len(self.media) else None
        d.mediumCount = len(self.media)
        d.cameraMedium = scene.desc.cameraMedium
        d.envSpectrum = int(env[0]["spectrum"]) if env else -1
        d.envScale = float(env[0]["scale"]) if env else 0.0
        cam = scene.camera
        c = d.camera
        c.cameraToWorld[:] = list(cam.cameraToWorld)
        c.tanHalfFov, c.aspect = cam.tanHalfFov, cam.width / cam.height
        c.lensRadius, c.focalDistance = cam.lensRadius, cam.focalDistance
        c.width, c.height, c.margin = cam.width, cam.height, 0
        d.filter = filter_v4(cam)
        d.sheenAlbedo = None if self.sheen is None else self.sheen.ctypes.data
        self.width, self.height = cam.width, cam.height

    def tabulated_scene(self, scene):
        """The extension scene with the same tabulated spectra (a table of equal samples is a
        constant, as a plain IOR is on the CPU), for the oracle."""
        rows = np.zeros(len(self.spectra), pa.SPECTRUM_DTYPE)
        knots = []
        for i, v in enumerate(self.spectra["v"]):
            rows[i]["scale"] = 1.0
            if np.all(v == v[0]):
                rows[i]["kind"], rows[i]["p"][0] = pa.SPEC_CONSTANT, v[0]
            else:
                rows[i]["kind"], rows[i]["offset"], rows[i]["count"] = pa.SPEC_TABLE95, len(knots), 95
                knots.extend(v.tolist())
        return pa.Scene(scene.prims, scene.materials, rows, np.asarray(knots, np.float32), scene.lights, scene.media,
                        scene.grid, scene.camera, scene.desc.cameraMedium, scene.defaults)


def params_v4(spp, first, seed=1, max_depth=-1, clamp=0.0):
    return pa.RenderParamsV4(spp, first, seed, max_depth, clamp)
