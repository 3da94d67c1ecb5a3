"""Edge cases of the render path on the MI355X against the oracle: a scene with no primitives
under a constant sky (every camera ray escapes), a scene with neither primitives nor lights (a
black film), and films of one pixel and of ragged sizes that fill no whole block of threads.

The reference renders such scenes without special cases (src/gpu/prender_gpu.cpp:1636-1731
uploads whatever counts the scene holds; src/integrators/path.cpp:140-153 adds the environment
for an escaped ray); here the empty geometry is a wide-BVH root with no children.
"""
import os

import numpy as np
import pytest

import oracle as orc
import prender_amd as pa
from test_pt_gpu import check_samples

pytestmark = pytest.mark.gpu

STUDIO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes", "studio", "scene.prscene.json")


def _without_geometry(s, keep_lights):
    lights = s.lights[s.lights["kind"] == pa.LIGHT_ENV] if keep_lights else s.lights[:0]
    cam = pa.CameraEx.from_buffer_copy(bytes(s.camera))
    cam.width, cam.height = 24, 16
    return pa.Scene(s.prims[:0], s.materials, s.spectra, s.knots, lights, None, None, cam, -1, s.defaults)


def test_empty_scene_under_a_constant_sky(plugin, host):
    s = _without_geometry(host.load_scene(STUDIO), keep_lights=True)
    assert len(s.prims) == 0 and (s.lights["kind"] == pa.LIGHT_ENV).sum() >= 1
    p = pa.render_params(spp=1, first_sample=0, seed=4, max_depth=8, rr_start=1)
    gpu, _ = check_samples("empty scene + sky", s, p, plugin, 1.0)
    assert np.all(np.any(gpu["L"] > 0, axis=1)), "every escaped camera ray sees the sky"
    p = pa.render_params(spp=3, first_sample=0, seed=4, max_depth=8, rr_start=1)
    g_sum, _ = plugin.render_ex(s, p)
    c_sum, _, _ = orc.OracleScene(s).render(p)
    np.testing.assert_allclose(g_sum, c_sum, rtol=1e-5, atol=1e-9)
    assert np.all(g_sum[..., 1] > 0)


def test_empty_scene_without_lights_is_black(plugin, host):
    s = _without_geometry(host.load_scene(STUDIO), keep_lights=False)
    assert len(s.prims) == 0 and len(s.lights) == 0
    for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
        p = pa.render_params(spp=2, first_sample=0, seed=4, max_depth=8, rr_start=1, light_sampler=sampler)
        g_sum, g_sq = plugin.render_ex(s, p)
        assert not g_sum.any() and not g_sq.any()
        assert not plugin.sample_radiance(s, p)["L"].any()


@pytest.mark.parametrize("size", [(1, 1), (7, 5), (67, 3)])
def test_one_pixel_and_ragged_films(plugin, host, size):
    w, h = size
    s = host.generate(1, w, h)
    assert (s.width, s.height) == (w, h)
    p = pa.render_params(spp=1, first_sample=3, seed=2, max_depth=8, rr_start=1)
    check_samples(f"c1 {w}x{h}", s, p, plugin, 0.95)
    p = pa.render_params(spp=4, first_sample=0, seed=2, max_depth=8, rr_start=3)
    g_sum, _ = plugin.render_ex(s, p)
    c_sum, _, _ = orc.OracleScene(s).render(p)
    assert g_sum.shape == (h, w, 3) and np.all(np.isfinite(g_sum))
    rel = np.abs(g_sum - c_sum) / np.maximum(np.abs(c_sum), 1e-3)
    assert np.mean(rel <= 1e-3) >= 0.9, f"{w}x{h}: films differ beyond diverged samples"
