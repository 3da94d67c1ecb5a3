"""Python binding of the MI355X prender plugin and its host library.

The plugin (libprender_gpu.so) exposes the reference's v4 C ABI (src/gpu/gpu_api.h) plus the
extension declared in include/prender_gpu.h; the host library (libprender_host.so) generates
the benchmark scenes and runs the progressive loop of the reference's GpuPathIntegrator
(src/integrators/gpu_path.cpp). This module mirrors those C structs with ctypes / numpy dtypes
and wraps the entry points. Loading fails loudly when a library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PLUGIN_PATH = os.path.join(HERE, "libprender_gpu.so")
HOST_PATH = os.path.join(HERE, "libprender_host.so")

API_VERSION = 4
EXT_VERSION = 4

# ---- enums (include/prender_gpu.h) ---------------------------------------------------------------
PRIM_TRIANGLE, PRIM_SPHERE = 0, 1
MAT_DIFFUSE, MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_BLACK, MAT_INTERFACE = 0, 1, 2, 3, 4
MAT_COATED_DIFFUSE, MAT_COATED_CONDUCTOR = 5, 6
SPEC_CONSTANT, SPEC_SIGMOID, SPEC_BLACKBODY, SPEC_PIECEWISE, SPEC_SELLMEIER, SPEC_CAUCHY, SPEC_TABLE95, SPEC_IMAGE = range(8)
LIGHT_AREA, LIGHT_POINT, LIGHT_DISTANT, LIGHT_ENV, LIGHT_SPOT, LIGHT_GONIO, LIGHT_SUN = 0, 1, 2, 3, 4, 5, 6
MEDIUM_HOMOGENEOUS, MEDIUM_GRID = 0, 1
FILTER_BOX, FILTER_GAUSSIAN, FILTER_BH = 0, 1, 2
LIGHT_SAMPLER_BVH, LIGHT_SAMPLER_POWER = 0, 1
RENDER_GUIDE, RENDER_GUIDE_TRAIN, RENDER_GUIDE_RESET = 1, 2, 4

# ---- numpy layouts of the POD arrays -------------------------------------------------------------
PRIM_DTYPE = np.dtype({
    "names": ["type", "material", "light", "mediumInside", "mediumOutside", "pad", "p0", "p1", "p2",
              "n0", "n1", "n2", "uv", "r", "flip"],
    "formats": ["<u4", "<i4", "<i4", "<i4", "<i4", "<u4", ("<f4", 3), ("<f4", 3), ("<f4", 3),
                ("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 6), "<f4", "<u4"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 36, 48, 60, 72, 84, 96, 120, 124],
    "itemsize": 128})
MATERIAL_DTYPE = np.dtype({
    "names": ["type", "albedo", "albedoB", "checkerScale", "eta", "k", "alphaX", "alphaY", "dispersive", "thin",
              "coatEta", "coatAlpha", "thickness", "coatG", "coatAlbedo", "coatDispersive", "coatMaxDepth",
              "coatSamples", "sheenColor", "sheenRoughness", "sheenWeight", "shading"],
    "formats": ["<u4", "<i4", "<i4", "<f4", "<i4", "<i4", "<f4", "<f4", "<u4", "<u4", "<i4", "<f4", "<f4", "<f4",
                "<i4", "<u4", "<i4", "<i4", "<i4", "<f4", "<f4", "<u4"],
    "itemsize": 88})
SPECTRUM_DTYPE = np.dtype({"names": ["kind", "scale", "offset", "count", "p"],
                           "formats": ["<u4", "<f4", "<i4", "<i4", ("<f4", 8)], "itemsize": 48})
LIGHT_DTYPE = np.dtype({"names": ["kind", "primitive", "spectrum", "scale", "position", "direction", "twoSided",
                                  "cosPower", "cosFalloffStart", "cosFalloffEnd", "profileOffset", "profileSize"],
                        "formats": ["<u4", "<i4", "<i4", "<f4", ("<f4", 3), ("<f4", 3), "<u4", "<f4", "<f4", "<f4",
                                    "<i4", "<i4"],
                        "offsets": [0, 4, 8, 12, 16, 28, 40, 44, 48, 52, 56, 60], "itemsize": 64})
MEDIUM_DTYPE = np.dtype({"names": ["kind", "sigmaA", "sigmaS", "scale", "g", "nx", "ny", "nz", "gridOffset",
                                   "maxDensity", "boundsMin", "boundsMax", "worldToMedium"],
                         "formats": ["<u4", "<i4", "<i4", "<f4", "<f4", "<i4", "<i4", "<i4", "<i8", "<f4",
                                     ("<f4", 3), ("<f4", 3), ("<f4", 12)],
                         "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 40, 44, 56, 68], "itemsize": 120})
RAY_DTYPE = np.dtype({"names": ["o", "d", "tMax"], "formats": [("<f4", 3), ("<f4", 3), "<f4"], "itemsize": 28})
HIT_DTYPE = np.dtype({"names": ["t", "prim", "b1", "b2"], "formats": ["<f4", "<i4", "<f4", "<f4"], "itemsize": 16})
SAMPLE_DTYPE = np.dtype({"names": ["L", "lam", "pdf", "pFilm", "weight", "flags"],
                         "formats": [("<f4", 4), ("<f4", 4), ("<f4", 4), ("<f4", 2), "<f4", "<u4"],
                         "itemsize": 64})


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class CameraEx(C.Structure):
    _fields_ = [("cameraToWorld", C.c_float * 12), ("tanHalfFov", C.c_float), ("lensRadius", C.c_float),
                ("focalDistance", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("filterType", C.c_int32), ("filterRadius", C.c_float), ("pad", C.c_int32)]


class SceneDescEx(C.Structure):
    _fields_ = [("apiVersion", C.c_int32), ("extVersion", C.c_int32),
                ("prims", C.c_void_p), ("primCount", C.c_int32),
                ("materials", C.c_void_p), ("materialCount", C.c_int32),
                ("spectra", C.c_void_p), ("spectrumCount", C.c_int32),
                ("knots", C.c_void_p), ("knotCount", C.c_int32),
                ("lights", C.c_void_p), ("lightCount", C.c_int32),
                ("media", C.c_void_p), ("mediumCount", C.c_int32),
                ("gridData", C.c_void_p), ("gridCount", C.c_int64),
                ("cameraMedium", C.c_int32), ("pad", C.c_int32),
                ("camera", CameraEx)]


class RenderParamsEx(C.Structure):
    _fields_ = [("spp", C.c_int32), ("firstSample", C.c_int32), ("seed", C.c_uint64), ("maxDepth", C.c_int32),
                ("clampLuminance", C.c_float), ("rrStartDepth", C.c_int32), ("russianRoulette", C.c_int32),
                ("lightSampler", C.c_int32), ("flags", C.c_int32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("globalMemory", C.c_uint64), ("computeUnits", C.c_int32),
                ("isGpu", C.c_int32), ("backend", C.c_char * 32)]


class ProfileStatsEx(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_int64 * 8), ("items", C.c_int64 * 8),
                ("closestRays", C.c_int64), ("closestNodes", C.c_int64), ("closestRecs", C.c_int64),
                ("shadowRays", C.c_int64), ("shadowNodes", C.c_int64), ("shadowRecs", C.c_int64),
                ("nodeBytes", C.c_int64), ("recBytes", C.c_int64)]


# ---- v4 structures (reference src/gpu/gpu_api.h:31-154) ------------------------------------
SPECTRUM95_DTYPE = np.dtype([("v", "<f4", 95)])
LIGHT_V4_DTYPE = np.dtype({"names": ["type", "primitive", "spectrum", "scale", "position", "twoSided", "cosPower",
                                     "area", "pmf", "pad"],
                           "formats": ["<u4", "<i4", "<i4", "<f4", ("<f4", 3), "<u4", "<f4", "<f4", "<f4", "<u4"],
                           "itemsize": 48})
MEDIUM_V4_DTYPE = np.dtype([("sigmaA", "<i4"), ("sigmaS", "<i4"), ("scale", "<f4"), ("g", "<f4")])
BVHNODE_V4_DTYPE = np.dtype({"names": ["bmin", "bmax", "offset", "nPrims", "axis", "pad"],
                             "formats": [("<f4", 3), ("<f4", 3), "<i4", "<u2", "u1", "u1"], "itemsize": 32})
V4_LIGHT_AREA, V4_LIGHT_POINT = 0, 1


class CameraV4(C.Structure):
    _fields_ = [("cameraToWorld", C.c_float * 12), ("tanHalfFov", C.c_float), ("aspect", C.c_float),
                ("lensRadius", C.c_float), ("focalDistance", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("margin", C.c_int32), ("pad", C.c_uint32)]


class FilterV4(C.Structure):
    _fields_ = [("type", C.c_uint32), ("radius", C.c_float), ("sigma", C.c_float), ("norm", C.c_float),
                ("cdf", C.c_float * 65)]


class SceneDescV4(C.Structure):
    _fields_ = [("apiVersion", C.c_int32), ("nodes", C.c_void_p), ("nodeCount", C.c_int32),
                ("primOrder", C.c_void_p), ("prims", C.c_void_p), ("primCount", C.c_int32),
                ("materials", C.c_void_p), ("materialCount", C.c_int32),
                ("lights", C.c_void_p), ("lightCount", C.c_int32),
                ("spectra", C.c_void_p), ("spectrumCount", C.c_int32),
                ("media", C.c_void_p), ("mediumCount", C.c_int32),
                ("cameraMedium", C.c_int32), ("envSpectrum", C.c_int32), ("envScale", C.c_float),
                ("envPmf", C.c_float), ("worldCenter", Float3), ("worldRadius", C.c_float),
                ("camera", CameraV4), ("filter", FilterV4), ("sheenAlbedo", C.c_void_p), ("cie", C.c_void_p),
                ("cieYIntegral", C.c_float)]


class RenderParamsV4(C.Structure):
    _fields_ = [("spp", C.c_int32), ("firstSample", C.c_int32), ("seed", C.c_uint64), ("maxDepth", C.c_int32),
                ("clampLuminance", C.c_float)]


assert C.sizeof(SceneDescV4) == 520 and C.sizeof(RenderParamsV4) == 24 and C.sizeof(CameraV4) == 80
assert C.sizeof(FilterV4) == 276 and LIGHT_V4_DTYPE.itemsize == 48 and SPECTRUM95_DTYPE.itemsize == 380

KERNEL_NAMES = ["camera", "intersect", "surface", "shadow", "finish", "gather", "guide", "material"]

assert C.sizeof(ProfileStatsEx) == 256
assert C.sizeof(SceneDescEx) == 208 and C.sizeof(RenderParamsEx) == 40 and C.sizeof(DeviceInfo) == 304


def render_params(spp=1, first_sample=0, seed=1, max_depth=-1, rr_start=1, rr=True, light_sampler=LIGHT_SAMPLER_BVH,
                  clamp=0.0, flags=0):
    return RenderParamsEx(spp, first_sample, seed, max_depth, clamp, rr_start, 1 if rr else 0, light_sampler, flags)


def flat_scene(npz, name):
    """A Scene from arrays saved under `name_*` keys (prims, materials, spectra, knots, lights, media,
    grid, camera bytes, camera_medium, max_depth, seed), as tests/golden/make_sample_scene.py writes them."""
    g = lambda f: npz[f"{name}_{f}"]  # noqa: E731
    cam = CameraEx.from_buffer_copy(g("camera").tobytes())
    defaults = {"spp": 16, "maxDepth": int(g("max_depth")), "rrStartDepth": 1, "seed": int(g("seed"))}
    return Scene(g("prims"), g("materials"), g("spectra"), g("knots"), g("lights"), g("media"), g("grid"), cam,
                 int(g("camera_medium")), defaults)


def measurement_rays(scene, n, seed):
    """Two deterministic ray batches over a scene, for traversal statistics: n primary rays through
    pixel-uniform film positions of the scene's camera (pinhole; the lens is ignored) and n rays
    from uniform points of the primitives' bounding box in uniform directions. float32 throughout."""
    rng = np.random.default_rng(seed)
    cam = scene.camera
    m = np.array(cam.cameraToWorld, np.float32).reshape(3, 4)
    w, h = float(cam.width), float(cam.height)
    fx = rng.uniform(0, w, n).astype(np.float32)
    fy = rng.uniform(0, h, n).astype(np.float32)
    t = np.float32(cam.tanHalfFov)
    aspect = np.float32(w / h)
    local = np.stack([(2 * fx / np.float32(w) - 1) * t * aspect, (1 - 2 * fy / np.float32(h)) * t,
                      np.ones(n, np.float32)], axis=1).astype(np.float32)
    d = local @ m[:, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cam_rays = np.zeros(n, RAY_DTYPE)
    cam_rays["o"] = m[:, 3]
    cam_rays["d"] = d.astype(np.float32)
    cam_rays["tMax"] = np.inf
    p = np.concatenate([scene.prims["p0"], scene.prims["p1"], scene.prims["p2"]])
    lo, hi = p.min(axis=0), p.max(axis=0)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rnd = np.zeros(n, RAY_DTYPE)
    rnd["o"], rnd["d"], rnd["tMax"] = o, v.astype(np.float32), np.inf
    return cam_rays, rnd


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


class Scene:
    """A flat scene: numpy arrays laid out as the C structs, plus the SceneDescEx pointing at them.

    Either filled by the host library's generators (``generate``) or assembled in Python for
    small test scenes. The arrays must stay alive while the descriptor is in use.
    """

    def __init__(self, prims, materials, spectra, knots, lights, media=None, grid=None, camera=None,
                 camera_medium=-1, defaults=None, handle=None, host=None):
        self.prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
        self.materials = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        self.spectra = np.ascontiguousarray(spectra, dtype=SPECTRUM_DTYPE)
        self.knots = np.ascontiguousarray(knots if knots is not None else np.zeros(0), dtype=np.float32)
        self.lights = np.ascontiguousarray(lights if lights is not None else np.zeros(0, LIGHT_DTYPE),
                                           dtype=LIGHT_DTYPE)
        self.media = np.ascontiguousarray(media if media is not None else np.zeros(0, MEDIUM_DTYPE),
                                          dtype=MEDIUM_DTYPE)
        self.grid = np.ascontiguousarray(grid if grid is not None else np.zeros(0), dtype=np.float32)
        self.camera = camera if camera is not None else CameraEx()
        self.defaults = defaults or {"spp": 16, "maxDepth": -1, "rrStartDepth": 1, "seed": 1}
        self._handle = handle
        self._host = host
        self.warnings, self.settings = [], {}
        self.desc = SceneDescEx()
        d = self.desc
        d.apiVersion, d.extVersion = API_VERSION, EXT_VERSION
        d.prims, d.primCount = _ptr(self.prims), len(self.prims)
        d.materials, d.materialCount = _ptr(self.materials), len(self.materials)
        d.spectra, d.spectrumCount = _ptr(self.spectra), len(self.spectra)
        d.knots, d.knotCount = _ptr(self.knots), len(self.knots)
        d.lights, d.lightCount = _ptr(self.lights), len(self.lights)
        d.media, d.mediumCount = _ptr(self.media), len(self.media)
        d.gridData, d.gridCount = _ptr(self.grid), len(self.grid)
        d.cameraMedium = camera_medium
        d.camera = self.camera

    @property
    def width(self):
        return self.camera.width

    @property
    def height(self):
        return self.camera.height

    def margin(self):
        r = self.camera.filterRadius if self.camera.filterType != FILTER_BOX else 0.5
        return int(np.ceil(r - 0.5))

    def desc_ptr(self):
        return C.byref(self.desc)

    def __del__(self):
        if self._handle is not None and self._host is not None:
            self._host.ph_scene_free(self._handle)
            self._handle = None


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_double)


class PhRenderControl(C.Structure):
    _fields_ = [("checkpointPath", C.c_char_p), ("resumePath", C.c_char_p), ("checkpointInterval", C.c_double),
                ("progressInterval", C.c_double), ("onProgress", PROGRESS_FN), ("user", C.c_void_p),
                ("cancel", C.POINTER(C.c_int32))]


def _load(path, what):
    if not os.path.exists(path):
        raise RuntimeError(f"{what} is not built: {path} is missing (run __graft_entry__.build())")
    return C.CDLL(path)


class Host:
    """libprender_host.so: scene generation, plugin loading, progressive render."""

    def __init__(self, path=HOST_PATH):
        L = self.lib = _load(path, "the prender host library")
        L.ph_generate.restype = C.c_void_p
        L.ph_generate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_char_p, C.c_int]
        L.ph_scene_free.argtypes = [C.c_void_p]
        L.ph_load_scene.restype = C.c_void_p
        L.ph_load_scene.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.ph_scene_warnings.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.ph_scene_file_settings.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int32)]
        L.ph_scene_desc.restype = C.POINTER(SceneDescEx)
        L.ph_scene_desc.argtypes = [C.c_void_p]
        L.ph_scene_defaults.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
        L.ph_plugin_open.restype = C.c_void_p
        L.ph_plugin_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.ph_plugin_close.argtypes = [C.c_void_p]
        L.ph_plugin_has_ext.argtypes = [C.c_void_p]
        L.ph_plugin_list_devices.argtypes = [C.c_void_p, C.POINTER(DeviceInfo), C.c_int]
        L.ph_render.argtypes = [C.c_void_p, C.c_int, C.POINTER(SceneDescEx), C.POINTER(RenderParamsEx), C.c_int,
                                C.c_double, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                C.c_char_p, C.c_int]
        L.ph_resolve_srgb.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_void_p]
        L.ph_fit_sigmoid.argtypes = [C.c_void_p, C.c_void_p]
        L.ph_rgb_fit_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ph_rgb_table_record.argtypes = [C.c_void_p, C.c_int]
        L.ph_preetham_sky.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_int, C.c_void_p]
        L.ph_envmap_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.ph_spectrum_value.restype = C.c_float
        L.ph_spectrum_value.argtypes = [C.c_void_p, C.c_void_p, C.c_float]
        L.ph_scene_fingerprint.restype = C.c_uint64
        L.ph_scene_fingerprint.argtypes = [C.POINTER(SceneDescEx)]
        L.ph_checkpoint_write.argtypes = [C.c_char_p, C.POINTER(SceneDescEx), C.c_int32, C.c_uint64, C.c_void_p,
                                          C.c_void_p, C.c_int32, C.c_char_p, C.c_int]
        L.ph_checkpoint_read.argtypes = [C.c_char_p, C.POINTER(SceneDescEx), C.c_int32, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_int32), C.c_char_p, C.c_int]
        L.ph_render_session.argtypes = [C.c_void_p, C.c_int, C.POINTER(SceneDescEx), C.POINTER(RenderParamsEx),
                                        C.c_int, C.c_double, C.POINTER(PhRenderControl), C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.ph_list_devices_json.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        L.ph_progress_json.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_char_p,
                                       C.c_int]
        L.ph_stage_json.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.ph_error_json.argtypes = [C.c_char_p, C.c_char_p, C.c_int]

    def generate(self, config, width=0, height=0, detail=1.0):
        err = C.create_string_buffer(512)
        h = self.lib.ph_generate(config, width, height, detail, err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        return self._take_scene(h)

    def load_scene(self, path, rigs=None, width=0, height=0):
        """Loads a .prscene.json file (ph_load_scene). rigs: None = the file's active rigs, else a
        list of rig names. Returns a Scene whose .warnings and .settings hold the loader's warnings and
        the file's render settings (clamp, time_limit, russian_roulette, light_bvh)."""
        err = C.create_string_buffer(2048)
        r = None if rigs is None else ",".join(rigs).encode()
        h = self.lib.ph_load_scene(str(path).encode(), r, width, height, err, 2048)
        if not h:
            raise RuntimeError(err.value.decode())
        n = self.lib.ph_scene_warnings(h, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.lib.ph_scene_warnings(h, buf, n + 1)
        clamp, limit, flags = C.c_float(), C.c_double(), C.c_int32()
        self.lib.ph_scene_file_settings(h, C.byref(clamp), C.byref(limit), C.byref(flags))
        s = self._take_scene(h)
        s.warnings = [w for w in buf.value.decode().split("\n") if w]
        s.settings = {"clamp": clamp.value, "time_limit": limit.value, "russian_roulette": bool(flags.value & 1),
                      "light_bvh": bool(flags.value & 2)}
        return s

    def _take_scene(self, h):
        d = self.lib.ph_scene_desc(h).contents

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * dt.itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dt).copy()

        vals = (C.c_int32 * 3)()
        seed = C.c_uint64()
        self.lib.ph_scene_defaults(h, vals, C.byref(seed))
        defaults = {"spp": vals[0], "maxDepth": vals[1], "rrStartDepth": vals[2], "seed": seed.value}
        cam = CameraEx.from_buffer_copy(d.camera)
        s = Scene(arr(d.prims, d.primCount, PRIM_DTYPE), arr(d.materials, d.materialCount, MATERIAL_DTYPE),
                  arr(d.spectra, d.spectrumCount, SPECTRUM_DTYPE), arr(d.knots, d.knotCount, np.dtype("<f4")),
                  arr(d.lights, d.lightCount, LIGHT_DTYPE), arr(d.media, d.mediumCount, MEDIUM_DTYPE),
                  arr(d.gridData, d.gridCount, np.dtype("<f4")), cam, d.cameraMedium, defaults)
        self.lib.ph_scene_free(h)
        return s

    def resolve_srgb(self, xyz, scale):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        out = np.zeros(xyz.shape, np.float32)
        self.lib.ph_resolve_srgb(xyz.ctypes.data, xyz.size // 3, scale, out.ctypes.data)
        return out

    def fit_sigmoid(self, rgb):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        out = np.zeros(3, np.float32)
        self.lib.ph_fit_sigmoid(rgb.ctypes.data, out.ctypes.data)
        return out

    def rgb_fit_tables(self):
        """The RGB-to-spectrum fit's inputs: (3, n) spectrum-to-linear-sRGB weights, the (3, 3)
        sRGB-to-XYZ matrix and its white."""
        n = self.lib.ph_rgb_fit_tables(None, None, None)
        w, m, white = np.zeros((3, n)), np.zeros((3, 3)), np.zeros(3)
        self.lib.ph_rgb_fit_tables(w.ctypes.data, m.ctypes.data, white.ctypes.data)
        return w, m, white

    def rgb_table_record(self):
        """The RGB-to-spectrum table record: [resolution, z nodes..., coefficients...]."""
        n = self.lib.ph_rgb_table_record(None, 0)
        out = np.zeros(n, np.float32)
        self.lib.ph_rgb_table_record(out.ctypes.data, n)
        return out

    def preetham_sky(self, sun_dir, turbidity, ground_albedo, width, height):
        """Preetham daylight map (height, width, 3) in linear sRGB, kcd/m^2, zenith row first."""
        d = np.ascontiguousarray(sun_dir, dtype=np.float32)
        out = np.zeros((height, width, 3), np.float32)
        if self.lib.ph_preetham_sky(d.ctypes.data, turbidity, ground_albedo, width, height, out.ctypes.data):
            raise ValueError("invalid sky arguments")
        return out

    def envmap_sample(self, rgb, u):
        """Samples of an equirectangular RGB map (height, width, 3) at pairs u (n, 2): rows of
        (u, v, pdf of the sample, density at (u, v)), and the map's mean luminance."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 4), np.float32)
        avg = C.c_float()
        if self.lib.ph_envmap_sample(rgb.ctypes.data, rgb.shape[1], rgb.shape[0], u.ctypes.data, len(u),
                                     out.ctypes.data, C.byref(avg)):
            raise ValueError("invalid environment map")
        return out, avg.value

    def spectrum_value(self, spec_row, knots, lam):
        s = np.ascontiguousarray(np.array([spec_row], dtype=SPECTRUM_DTYPE))
        k = np.ascontiguousarray(knots, dtype=np.float32)
        return self.lib.ph_spectrum_value(s.ctypes.data, k.ctypes.data if len(k) else None, lam)

    def fingerprint(self, scene):
        return self.lib.ph_scene_fingerprint(scene.desc_ptr())

    def checkpoint_write(self, path, scene, max_depth, seed, film_sum, film_sq, passes):
        s = np.ascontiguousarray(film_sum, dtype=np.float64)
        q = np.ascontiguousarray(film_sq, dtype=np.float64) if film_sq is not None else None
        err = C.create_string_buffer(1024)
        if self.lib.ph_checkpoint_write(str(path).encode(), scene.desc_ptr(), max_depth, seed, s.ctypes.data,
                                        q.ctypes.data if q is not None else None, passes, err, 1024) != 0:
            raise RuntimeError(err.value.decode())

    def checkpoint_read(self, path, scene, max_depth, seed, variance=True):
        shape = (scene.height, scene.width, 3)
        s = np.zeros(shape, np.float64)
        q = np.zeros(shape, np.float64) if variance else None
        passes = C.c_int32()
        err = C.create_string_buffer(1024)
        if self.lib.ph_checkpoint_read(str(path).encode(), scene.desc_ptr(), max_depth, seed, s.ctypes.data,
                                       q.ctypes.data if q is not None else None, C.byref(passes), err, 1024) != 0:
            raise RuntimeError(err.value.decode())
        return s, q, passes.value

    def open_plugin(self, path=PLUGIN_PATH):
        err = C.create_string_buffer(512)
        h = self.lib.ph_plugin_open(path.encode(), err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        return h

    def render(self, plugin, scene, params, spp, device=0, time_limit=0.0, variance=True):
        n = scene.width * scene.height * 3
        s = np.zeros(n, np.float64)
        sq = np.zeros(n, np.float64) if variance else None
        done = C.c_int32()
        secs = C.c_double()
        err = C.create_string_buffer(1024)
        rc = self.lib.ph_render(plugin, device, scene.desc_ptr(), C.byref(params), spp, time_limit, s.ctypes.data,
                                sq.ctypes.data if sq is not None else None, C.byref(done), C.byref(secs), err, 1024)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return s.reshape(scene.height, scene.width, 3), (sq.reshape(scene.height, scene.width, 3)
                                                          if sq is not None else None), done.value, secs.value

    def render_session(self, plugin, scene, params, spp, device=0, time_limit=0.0, checkpoint=None, resume=None,
                       checkpoint_interval=60.0, progress_interval=0.5, on_progress=None, cancel=None,
                       variance=True):
        """Progressive render with resume, checkpoints and progress (ph_render_session).

        on_progress(stage, fraction, elapsed, spp) may return True to cancel; cancel is an optional
        ctypes.c_int32 the caller may set from another thread. Returns (sum, sumsq, spp_done, scale).
        """
        shape = (scene.height, scene.width, 3)
        s = np.zeros(shape, np.float64)
        sq = np.zeros(shape, np.float64) if variance else None
        cb = PROGRESS_FN(0)
        if on_progress is not None:
            def _cb(_user, stage, fraction, elapsed, spp_done):
                return 1 if on_progress(stage.decode(), fraction, elapsed, spp_done) else 0
            cb = PROGRESS_FN(_cb)
        ctl = PhRenderControl(str(checkpoint).encode() if checkpoint else None,
                              str(resume).encode() if resume else None, checkpoint_interval, progress_interval, cb,
                              None, C.pointer(cancel) if cancel is not None else None)
        done = C.c_int32()
        scale = C.c_double()
        err = C.create_string_buffer(1024)
        rc = self.lib.ph_render_session(plugin, device, scene.desc_ptr(), C.byref(params), spp, time_limit,
                                        C.byref(ctl), s.ctypes.data, sq.ctypes.data if sq is not None else None,
                                        C.byref(done), C.byref(scale), err, 1024)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return s, sq, done.value, scale.value

    @staticmethod
    def _line(fn, *args):
        n = fn(*args, None, 0)
        buf = C.create_string_buffer(n + 1)
        fn(*args, buf, n + 1)
        return buf.value.decode()

    def list_devices_json(self, plugin_path=PLUGIN_PATH, cpu_threads=0):
        return self._line(self.lib.ph_list_devices_json, plugin_path.encode() if plugin_path else None, cpu_threads)

    def progress_json(self, stage, fraction, elapsed, spp, accept_rate=-1.0):
        return self._line(self.lib.ph_progress_json, stage.encode(), fraction, elapsed, spp, accept_rate)

    def stage_json(self, name):
        return self._line(self.lib.ph_stage_json, name.encode())

    def error_json(self, message):
        return self._line(self.lib.ph_error_json, message.encode())


class Plugin:
    """Direct binding of libprender_gpu.so's C ABI."""

    def __init__(self, path=PLUGIN_PATH):
        L = self.lib = _load(path, "the MI355X plugin")
        L.prgpu_api_version.restype = C.c_int
        L.prgpu_api_ext_version.restype = C.c_int
        L.prgpu_list_devices.argtypes = [C.POINTER(DeviceInfo), C.c_int]
        L.prgpu_trace_rays.argtypes = [C.c_int, C.POINTER(SceneDescEx), C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                       C.c_char_p, C.c_int]
        L.prgpu_render_ex.argtypes = [C.c_int, C.POINTER(SceneDescEx), C.POINTER(RenderParamsEx), C.c_void_p,
                                      C.c_void_p, C.c_char_p, C.c_int]
        L.prgpu_sample_radiance.argtypes = [C.c_int, C.POINTER(SceneDescEx), C.POINTER(RenderParamsEx), C.c_void_p,
                                            C.c_char_p, C.c_int]
        L.prgpu_render.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.prgpu_release.argtypes = [C.c_int]
        L.prgpu_film_reset.argtypes = [C.c_int, C.c_int, C.c_int]
        L.prgpu_film_read.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.prgpu_profile.argtypes = [C.c_int, C.c_int]
        L.prgpu_profile_read.argtypes = [C.c_int, C.POINTER(ProfileStatsEx), C.c_int]
        L.prgpu_guide_param_count.restype = C.c_int
        L.prgpu_guide_params.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int]
        L.prgpu_guide_eval.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.prgpu_guide_train.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p]
        L.prgpu_guide_stats.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_int)]

    def version(self):
        return self.lib.prgpu_api_version(), self.lib.prgpu_api_ext_version()

    def devices(self):
        arr = (DeviceInfo * 16)()
        n = self.lib.prgpu_list_devices(arr, 16)
        return [arr[i] for i in range(n)]

    @staticmethod
    def _check(rc, err):
        if rc != 0:
            raise RuntimeError(err.value.decode())

    def trace(self, scene, rays, any_hit=False, device=0):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        err = C.create_string_buffer(1024)
        self._check(self.lib.prgpu_trace_rays(device, scene.desc_ptr(), rays.ctypes.data if len(rays) else None,
                                              len(rays), 1 if any_hit else 0,
                                              hits.ctypes.data if len(hits) else None, err, 1024), err)
        return hits

    TUNE_KEYS = ("trace_blocks_per_cu", "shade_blocks_per_cu", "path_budget", "guide_batch", "guide_steps",
                 "guide_min_steps", "guide_lr", "guide_bounces", "guide_blend", "graph_rounds",
                 "guide_replay", "material_sort", "ray_sort")

    def tune_set(self, key, value, device=0):
        self.lib.prgpu_tune_set.argtypes = [C.c_int, C.c_char_p, C.c_double]
        if self.lib.prgpu_tune_set(device, key.encode(), float(value)) != 0:
            raise ValueError(f"tuning key {key!r}: unknown or value {value!r} out of range")

    def tune_get(self, key, device=0):
        self.lib.prgpu_tune_get.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_double)]
        v = C.c_double()
        if self.lib.prgpu_tune_get(device, key.encode(), C.byref(v)) != 0:
            raise ValueError(f"unknown tuning key {key!r}")
        return v.value

    def tuning(self, device=0):
        """The device's run-time tuning, and the tuning file it started from (None if none)."""
        self.lib.prgpu_tune_source.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
        path, err = C.create_string_buffer(1024), C.create_string_buffer(1024)
        found = self.lib.prgpu_tune_source(path, 1024, err, 1024) == 0
        try:
            values = {k: self.tune_get(k, device) for k in self.TUNE_KEYS}
        except ValueError:  # no such device
            values = None
        return {"values": values, "file": path.value.decode() if found else None,
                "file_error": err.value.decode() or None}

    def trace_noalloc(self, scene, rays_dev, n, hits_dev, any_hit=False, device=0):
        """prgpu_trace_rays_noalloc: rays_dev / hits_dev are device addresses (ints) of this GPU's
        memory holding n RayEx / HitEx records; the scene's geometry must already be resident."""
        f = self.lib.prgpu_trace_rays_noalloc
        f.argtypes = [C.c_int, C.POINTER(SceneDescEx), C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_char_p, C.c_int]
        err = C.create_string_buffer(1024)
        self._check(f(device, scene.desc_ptr(), rays_dev, n, 1 if any_hit else 0, hits_dev, err, 1024), err)

    def render_ex(self, scene, params, device=0, variance=True):
        n = scene.width * scene.height * 3
        s = np.zeros(n, np.float64)
        sq = np.zeros(n, np.float64) if variance else None
        err = C.create_string_buffer(1024)
        self._check(self.lib.prgpu_render_ex(device, scene.desc_ptr(), C.byref(params), s.ctypes.data,
                                             sq.ctypes.data if sq is not None else None, err, 1024), err)
        shape = (scene.height, scene.width, 3)
        return s.reshape(shape), (sq.reshape(shape) if sq is not None else None)

    def sample_radiance(self, scene, params, device=0):
        m = scene.margin()
        n = (scene.width + 2 * m) * (scene.height + 2 * m)
        out = np.zeros(n, SAMPLE_DTYPE)
        err = C.create_string_buffer(1024)
        self._check(self.lib.prgpu_sample_radiance(device, scene.desc_ptr(), C.byref(params), out.ctypes.data, err,
                                                   1024), err)
        return out

    def render_v4(self, desc, params, width, height, out=None, device=0):
        """The reference's v4 entry point: ADDS this call's XYZ sums into `out` (float32, h x w x 3)."""
        if out is None:
            out = np.zeros((height, width, 3), np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous and out.size == width * height * 3
        err = C.create_string_buffer(1024)
        self._check(self.lib.prgpu_render(device, C.byref(desc), C.byref(params), out.ctypes.data, err, 1024), err)
        return out

    def release(self, device=0):
        return self.lib.prgpu_release(device)

    # ---- path-guiding model ------------------------------------------------------------------
    def guide_params(self, params=None, device=0):
        """Reads the device's model parameters, or replaces them (and resets the optimiser)."""
        n = self.lib.prgpu_guide_param_count()
        if params is None:
            out = np.zeros(n, np.float32)
            if self.lib.prgpu_guide_params(device, out.ctypes.data, n, 0) != 0:
                raise RuntimeError("prgpu_guide_params failed")
            return out
        P = np.ascontiguousarray(params, dtype=np.float32)
        if self.lib.prgpu_guide_params(device, P.ctypes.data, len(P), 1) != 0:
            raise RuntimeError("prgpu_guide_params failed")
        return P

    def guide_eval(self, recs9, device=0):
        R = np.ascontiguousarray(recs9, dtype=np.float32).reshape(-1, 9)
        out = np.zeros((len(R), 41), np.float32)
        if self.lib.prgpu_guide_eval(device, R.ctypes.data, len(R), out.ctypes.data) != 0:
            raise RuntimeError("prgpu_guide_eval failed")
        return out

    def guide_grad(self, recs16, device=0):
        R = np.ascontiguousarray(recs16, dtype=np.float32).reshape(-1, 16)
        g = np.zeros(self.lib.prgpu_guide_param_count(), np.float32)
        st = np.zeros(2, np.float32)
        if self.lib.prgpu_guide_train(device, R.ctypes.data, len(R), 0, 0.0, g.ctypes.data, st.ctypes.data) != 0:
            raise RuntimeError("prgpu_guide_train failed")
        return g, st

    def guide_records(self, max_records, device=0):
        """The training records the last training batch kept (n x 16 float32), and how many it produced."""
        f = self.lib.prgpu_guide_records
        f.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        out = np.zeros((max_records, 16), np.float32)
        produced = C.c_int()
        n = f(device, out.ctypes.data, max_records, C.byref(produced))
        if n < 0:
            raise RuntimeError("prgpu_guide_records failed")
        return out[:n], produced.value

    def guide_stats(self, device=0):
        """(summed weight, summed weighted loss) of the model's last step, and its step count."""
        st = np.zeros(2, np.float32)
        steps = C.c_int()
        if self.lib.prgpu_guide_stats(device, st.ctypes.data, C.byref(steps)) != 0:
            return None
        return float(st[0]), float(st[1]), steps.value

    def guide_train(self, recs16, steps, lr=1e-3, device=0):
        R = np.ascontiguousarray(recs16, dtype=np.float32).reshape(-1, 16)
        st = np.zeros(2, np.float32)
        if self.lib.prgpu_guide_train(device, R.ctypes.data, len(R), steps, lr, None, st.ctypes.data) != 0:
            raise RuntimeError("prgpu_guide_train failed")
        return st

    # ---- device-resident film and profiling --------------------------------------------------
    def film_reset(self, width, height, device=0):
        if self.lib.prgpu_film_reset(device, width, height) != 0:
            raise RuntimeError("prgpu_film_reset failed")

    def render_resident(self, scene, params, device=0):
        """Renders into the device-resident film (no host transfer)."""
        err = C.create_string_buffer(1024)
        self._check(self.lib.prgpu_render_ex(device, scene.desc_ptr(), C.byref(params), None, None, err, 1024), err)

    def film_read(self, width, height, device=0, variance=True):
        s = np.zeros(width * height * 3, np.float64)
        sq = np.zeros_like(s) if variance else None
        if self.lib.prgpu_film_read(device, s.ctypes.data, sq.ctypes.data if sq is not None else None) != 0:
            raise RuntimeError("prgpu_film_read failed")
        return s.reshape(height, width, 3), (sq.reshape(height, width, 3) if sq is not None else None)

    def profile(self, mode, device=0):
        if self.lib.prgpu_profile(device, mode) != 0:
            raise RuntimeError("prgpu_profile failed")

    def profile_read(self, reset=True, device=0):
        st = ProfileStatsEx()
        if self.lib.prgpu_profile_read(device, C.byref(st), 1 if reset else 0) != 0:
            raise RuntimeError("prgpu_profile_read failed")
        out = {"kernels": {}}
        for i, name in enumerate(KERNEL_NAMES):
            out["kernels"][name] = {"ms": st.ms[i], "launches": st.launches[i], "items": st.items[i]}
        for f in ("closestRays", "closestNodes", "closestRecs", "shadowRays", "shadowNodes", "shadowRecs", "nodeBytes",
                  "recBytes"):
            out[f] = getattr(st, f)
        return out
