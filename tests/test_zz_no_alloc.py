"""The allocation-free ray query (prgpu_trace_rays_noalloc) against prgpu_trace_rays.

Device memory for the rays and hits comes from hipMalloc of the same HIP runtime the plugin
links (/opt/rocm's libamdhip64), through ctypes. Two calls in a row on configuration 2 (the deep
2M-triangle BVH, whose traversal stacks spill into the scratch), the second with a smaller batch,
then any-hit queries: ids, distances and barycentrics must equal prgpu_trace_rays' bit for bit.
A third case checks the errors: geometry that is not resident, and host pointers instead of
device memory, are refused without a launch.
"""
import ctypes as C

import numpy as np
import pytest

import prender_amd as pa
import raygen

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipSetDevice.argtypes = [C.c_int]
    return lib


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipSetDevice(0) == 0
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 16)) == 0, "hipMalloc failed"

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(plugin, hip, scene, rays, any_hit=False):
    n = len(rays)
    dr, dh = DevArray(hip, rays.nbytes), DevArray(hip, n * pa.HIT_DTYPE.itemsize)
    try:
        dr.put(rays)
        # poison the output so that a hit the query does not write is seen
        dh.put(np.full(n, 0xFF, np.uint8).repeat(pa.HIT_DTYPE.itemsize))
        plugin.trace_noalloc(scene, dr.p.value, n, dh.p.value, any_hit=any_hit)
        return dh.get(np.zeros(n, pa.HIT_DTYPE))
    finally:
        dr.free()
        dh.free()


def _same(a, b):
    for f in ("prim", "t", "b1", "b2"):
        assert np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)), f"{f} differs"


def test_noalloc_matches_trace_rays(plugin, host, hip):
    s = host.generate(2, 480, 270)
    lo, hi = raygen.scene_bounds(s)
    rays = raygen.random_rays(lo, hi, 200000, 11)
    want = plugin.trace(s, rays)  # also makes the geometry resident
    got = _noalloc(plugin, hip, s, rays)
    _same(got, want)
    small = rays[:777].copy()
    got2 = _noalloc(plugin, hip, s, small)
    _same(got2, want[:777])
    print(f"noalloc: {len(rays)} then {len(small)} rays equal prgpu_trace_rays; hit rate {np.mean(got['prim'] >= 0):.3f}",
          flush=True)
    shadow = rays[:50000].copy()
    shadow["tMax"] = np.random.default_rng(5).uniform(0.1, 20.0, len(shadow)).astype(np.float32)
    want_any = plugin.trace(s, shadow, any_hit=True)
    got_any = _noalloc(plugin, hip, s, shadow, any_hit=True)
    assert np.array_equal(got_any["prim"] >= 0, want_any["prim"] >= 0)


def test_noalloc_refuses_what_it_cannot_serve(plugin, host, hip):
    s1 = host.generate(1, 64, 64)
    s2 = host.generate(2, 160, 90)
    lo, hi = raygen.scene_bounds(s1)
    rays = raygen.random_rays(lo, hi, 1000, 3)
    plugin.trace(s2, rays[:16])  # configuration 2 resident: configuration 1's geometry is not
    dr, dh = DevArray(hip, rays.nbytes), DevArray(hip, len(rays) * pa.HIT_DTYPE.itemsize)
    try:
        dr.put(rays)
        with pytest.raises(RuntimeError, match="not resident"):
            plugin.trace_noalloc(s1, dr.p.value, len(rays), dh.p.value)
        plugin.trace(s1, rays[:16])
        host_hits = np.zeros(len(rays), pa.HIT_DTYPE)
        with pytest.raises(RuntimeError, match="device allocations"):
            plugin.trace_noalloc(s1, dr.p.value, len(rays), host_hits.ctypes.data)
        with pytest.raises(RuntimeError, match="device allocations"):
            plugin.trace_noalloc(s1, dr.p.value, len(rays) + 1, dh.p.value)
        plugin.trace_noalloc(s1, dr.p.value, len(rays), dh.p.value)
        got = dh.get(np.zeros(len(rays), pa.HIT_DTYPE))
        _same(got, plugin.trace(s1, rays))
    finally:
        dr.free()
        dh.free()
