"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of the reference CPU path tracer (src/integrators/path.cpp and
everything under it). Only tests/, the cpu_baseline leg of bench.py and __graft_entry__.smoke()
use it, as the checker. Struct layouts come from the product binding (prender_amd), which
mirrors include/prender_gpu.h.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "ar-renderer_amd"))
import prender_amd as pa  # noqa: E402

LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, fp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int32)
        L.or_scene_create.restype = vp
        L.or_scene_create.argtypes = [C.POINTER(pa.SceneDescEx)]
        L.or_scene_destroy.argtypes = [vp]
        L.or_scene_info.argtypes = [vp, vp]
        L.or_world_sphere.argtypes = [vp, vp]
        L.or_bvh_dump.argtypes = [vp, vp, vp]
        L.or_trace.argtypes = [vp, vp, C.c_int64, C.c_int, vp, C.c_int]
        L.or_trace_brute.argtypes = [vp, vp, C.c_int64, vp, vp, C.c_int]
        L.or_light_bvh_dump.argtypes = [vp, vp, vp]
        L.or_light_sample.argtypes = [vp, C.c_int, vp, vp, C.c_float, fp]
        L.or_light_pmf.restype = C.c_float
        L.or_light_pmf.argtypes = [vp, C.c_int, vp, vp, C.c_int]
        L.or_camera_ray.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, vp]
        L.or_sample_radiance.argtypes = [vp, C.POINTER(pa.RenderParamsEx), vp, C.c_int]
        L.or_render.restype = C.c_double
        L.or_render.argtypes = [vp, C.POINTER(pa.RenderParamsEx), vp, vp, C.c_int]
        L.or_rng_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, vp]
        L.or_sample_visible.argtypes = [C.c_float, vp, vp]
        L.or_cie.argtypes = [C.c_float, vp]
        L.or_cie_y_integral.restype = C.c_float
        L.or_to_xyz.argtypes = [vp, vp, vp, vp]
        for f in ("or_spectrum_eval",):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [vp, vp, C.c_int, C.c_float]
        for f in ("or_spectrum_max", "or_spectrum_y"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [vp, vp, C.c_int]
        L.or_bxdf.argtypes = [vp, C.c_int, vp, vp, C.c_float, vp, vp]
        L.or_closure.argtypes = [vp, C.c_int, vp, vp, C.c_float, vp, C.c_int, C.c_int, vp]
        L.or_sheen_albedo.restype = C.c_float
        L.or_sheen_albedo.argtypes = [C.c_float, C.c_float]
        L.or_sheen_table.argtypes = [vp]
        for f in ("or_fr_dielectric", "or_fr_complex", "or_hg"):
            getattr(L, f).restype = C.c_float
        L.or_fr_dielectric.argtypes = [C.c_float, C.c_float]
        L.or_fr_complex.argtypes = [C.c_float, C.c_float, C.c_float]
        L.or_hg.argtypes = [C.c_float, C.c_float]
        L.or_sample_hg.argtypes = [vp, C.c_float, vp, vp, vp]
        L.or_filter.argtypes = [C.c_int, C.c_float, vp, C.c_int, vp, vp]
        L.or_light_sample_li.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.or_tex_lookup.argtypes = [vp, vp, C.c_int, vp]
        L.or_perturb_shading.argtypes = [vp, vp, vp, C.c_int, vp]
        L.or_light_pdf_li.restype = C.c_float
        L.or_light_pdf_li.argtypes = [vp, C.c_int, vp, vp]
        L.or_light_power.restype = C.c_float
        L.or_light_power.argtypes = [vp, C.c_int]
        L.or_medium_tr.argtypes = [vp, C.c_int, vp, vp, C.c_float, vp, vp, vp]
        L.or_medium_sample.argtypes = [vp, C.c_int, vp, vp, C.c_float, C.c_float, vp, vp, vp]
        L.or_guide_param_count.restype = C.c_int
        L.or_guide_forward.argtypes = [vp, vp, C.c_int, vp]
        L.or_guide_mix_pdf.restype = C.c_double
        L.or_guide_mix_pdf.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.or_guide_mix_sample.argtypes = [vp, C.c_float, C.c_float, C.c_float, vp]
        L.or_guide_loss_grad.restype = C.c_double
        L.or_guide_loss_grad.argtypes = [vp, vp, C.c_int, vp]
        _lib = L
    return _lib


# ---- closures (plain, coated, sheen) ---------------------------------------------------------------
CLOSURE_DESC = 32  # floats per closure descriptor (layout in pt_oracle.cpp, or_closure)


def closure_desc(kind, alpha=(0.0, 0.0), eta=1.5, R=(0.5, 0.5, 0.5, 0.5), ceta=(0.2, 0.9, 1.1, 1.3),
                 k=(3.9, 2.4, 2.2, 1.9), coat_alpha=0.0, thickness=0.01, g=0.0, albedo=(0, 0, 0, 0), max_depth=64,
                 n_samples=1, sheen=0.0, sheen_alpha=0.3, sheen_color=(1, 1, 1, 1)):
    """Closure descriptor: kind 1 diffuse, 2 conductor, 3 dielectric, 4 thin, 5 coated diffuse,
    6 coated conductor; eta is the dielectric's or the coat's IOR."""
    return [kind, alpha[0], alpha[1], eta, *R, *ceta, *k, coat_alpha, thickness, g, *albedo, max_depth, n_samples, 1,
            sheen, sheen_alpha, *sheen_color]


def closure(desc, op, wo, wi=(0, 0, 1), uc=0.0, u=(0, 0), mode=0, rt=3):
    """One call of the closure probe: op 0 f, 1 pdf, 2 sample, 3 (flags, exact)."""
    d = np.ascontiguousarray(desc, dtype=np.float32)
    a = np.ascontiguousarray(wo, dtype=np.float32)
    b = np.ascontiguousarray(wi, dtype=np.float32)
    uu = np.ascontiguousarray(u, dtype=np.float32)
    out = np.zeros(12, np.float32)
    lib().or_closure(d.ctypes.data, op, a.ctypes.data, b.ctypes.data, float(uc), uu.ctypes.data, mode, rt,
                     out.ctypes.data)
    return out


def sheen_table():
    t = np.zeros(16 * 32, np.float32)
    lib().or_sheen_table(t.ctypes.data)
    return t


# ---- path-guiding model (FP64) -------------------------------------------------------------------
def guide_param_count():
    return lib().or_guide_param_count()


def guide_forward(params, recs9):
    P = np.ascontiguousarray(params, dtype=np.float32)
    R = np.ascontiguousarray(recs9, dtype=np.float32).reshape(-1, 9)
    out = np.zeros((len(R), 41), np.float32)
    lib().or_guide_forward(P.ctypes.data, R.ctypes.data, len(R), out.ctypes.data)
    return out


def guide_mix_pdf(out41, w):
    o = np.ascontiguousarray(out41, dtype=np.float32)
    wv = np.ascontiguousarray(w, dtype=np.float32)
    sel = C.c_double()
    p = lib().or_guide_mix_pdf(o.ctypes.data, wv.ctypes.data, C.byref(sel))
    return p, sel.value


def guide_mix_sample(out41, ul, u1, u2):
    o = np.ascontiguousarray(out41, dtype=np.float32)
    w = np.zeros(3, np.float32)
    lib().or_guide_mix_sample(o.ctypes.data, ul, u1, u2, w.ctypes.data)
    return w


def guide_set_blend(b):
    """Weight of the MIS-proposal term in the training loss (the rest is the mixture's own term)."""
    lib().or_guide_set_blend.argtypes = [C.c_double]
    lib().or_guide_set_blend(float(b))


def guide_set_sel_min(s):
    """Floor of the selection probability s (its ceiling is 0.9)."""
    lib().or_guide_set_sel_min.argtypes = [C.c_double]
    lib().or_guide_set_sel_min(float(s))


def guide_set_encoding(e):
    lib().or_guide_set_encoding.argtypes = [C.c_int]
    lib().or_guide_set_encoding(int(e))


def guide_loss_grad(params, recs16):
    P = np.ascontiguousarray(params, dtype=np.float32)
    R = np.ascontiguousarray(recs16, dtype=np.float32).reshape(-1, 16)
    g = np.zeros(len(P), np.float32)
    loss = lib().or_guide_loss_grad(P.ctypes.data, R.ctypes.data, len(R), g.ctypes.data)
    return loss, g


def f32(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float32).ravel())


class OracleScene:
    """The oracle's view of a prender_amd.Scene (its own binary BVH and light BVH)."""

    def __init__(self, scene):
        self.scene = scene
        self.h = lib().or_scene_create(C.byref(scene.desc))
        if not self.h:
            raise RuntimeError("oracle scene creation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    def info(self):
        out = np.zeros(4, np.int64)
        lib().or_scene_info(self.h, out.ctypes.data)
        return {"bvh_nodes": int(out[0]), "light_bvh_nodes": int(out[1]), "infinite_lights": int(out[2]),
                "margin": int(out[3])}

    def world_sphere(self):
        out = np.zeros(4, np.float32)
        lib().or_world_sphere(self.h, out.ctypes.data)
        return out

    def trace(self, rays, any_hit=False, threads=0):
        rays = np.ascontiguousarray(rays, dtype=pa.RAY_DTYPE)
        hits = np.zeros(len(rays), pa.HIT_DTYPE)
        lib().or_trace(self.h, rays.ctypes.data, len(rays), 1 if any_hit else 0, hits.ctypes.data,
                       threads or os.cpu_count())
        return hits

    def trace_brute(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, dtype=pa.RAY_DTYPE)
        hits = np.zeros(len(rays), pa.HIT_DTYPE)
        ties = np.zeros(len(rays), np.int32)
        lib().or_trace_brute(self.h, rays.ctypes.data, len(rays), hits.ctypes.data, ties.ctypes.data,
                             threads or os.cpu_count())
        return hits, ties

    def light_bvh(self):
        n = self.info()["light_bvh_nodes"]
        f = np.zeros((max(n, 1), 12), np.float32)
        iv = np.zeros((max(n, 1), 3), np.int32)
        lib().or_light_bvh_dump(self.h, f.ctypes.data, iv.ctypes.data)
        return f[:n], iv[:n]

    def light_sample(self, p, n, u, bvh=True):
        pmf = C.c_float()
        pv, nv = f32(*p), f32(*n)  # named: the arrays must outlive the call that reads their addresses
        li = lib().or_light_sample(self.h, 1 if bvh else 0, pv.ctypes.data, nv.ctypes.data, u, C.byref(pmf))
        return li, pmf.value

    def light_pmf(self, p, n, light, bvh=True):
        pv, nv = f32(*p), f32(*n)
        return lib().or_light_pmf(self.h, 1 if bvh else 0, pv.ctypes.data, nv.ctypes.data, light)

    def camera_ray(self, fx, fy, lx=0.5, ly=0.5):
        out = np.zeros(6, np.float32)
        lib().or_camera_ray(self.h, fx, fy, lx, ly, out.ctypes.data)
        return out

    def sample_radiance(self, params, threads=0):
        m = self.scene.margin()
        n = (self.scene.width + 2 * m) * (self.scene.height + 2 * m)
        out = np.zeros(n, pa.SAMPLE_DTYPE)
        lib().or_sample_radiance(self.h, C.byref(params), out.ctypes.data, threads or os.cpu_count())
        return out

    def render(self, params, variance=True, threads=0):
        n = self.scene.width * self.scene.height * 3
        s = np.zeros(n, np.float64)
        sq = np.zeros(n, np.float64) if variance else None
        secs = lib().or_render(self.h, C.byref(params), s.ctypes.data, sq.ctypes.data if sq is not None else None,
                               threads or os.cpu_count())
        shape = (self.scene.height, self.scene.width, 3)
        return s.reshape(shape), (sq.reshape(shape) if sq is not None else None), secs
