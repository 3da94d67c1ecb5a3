"""The reference's v4 entry point prgpu_render (src/gpu/gpu_api.h:168-172) on the MI355X plugin.

The v4 scene is built as the reference exporter builds it (tests/v4scene.py). Its output must
equal, sample for sample, the oracle's film for the extension scene with the same tabulated
spectra: both draw the CPU path tracer's sample streams, so the films differ only where a
sample's path diverged in the last ulp (see test_pt_gpu.py).
"""
import numpy as np
import pytest

import oracle as orc
import prender_amd as pa
from v4scene import V4Scene, params_v4

pytestmark = pytest.mark.gpu


def test_v4_render_matches_oracle_same_samples(plugin, host):
    s = host.generate(1, 64, 64)
    v4 = V4Scene(host, s)
    out = plugin.render_v4(v4.desc, params_v4(4, 0, seed=1, max_depth=8), v4.width, v4.height)
    ref, _, _ = orc.OracleScene(v4.tabulated_scene(s)).render(
        pa.render_params(spp=4, first_sample=0, seed=1, max_depth=8, rr_start=1), variance=False)
    rel = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-3)
    close = np.mean(rel <= 1e-3)
    print(f"v4 film vs oracle: fraction within 1e-3 = {close:.4f}, sums {out.sum():.6e} {ref.sum():.6e}", flush=True)
    assert close >= 0.97
    np.testing.assert_allclose(out.sum(), ref.sum(), rtol=5e-3)


def test_v4_render_uploads_its_own_geometry(plugin, host):
    """prgpu_render after a render of other geometry traces its own scene's primitives: the same
    film as after a render of its own geometry."""
    s = host.generate(1, 48, 48)
    v4 = V4Scene(host, s)
    first = plugin.render_v4(v4.desc, params_v4(2, 0, seed=3, max_depth=8), v4.width, v4.height)
    other = host.generate(6, 32, 32, 0.25)  # different triangles
    plugin.render_ex(other, pa.render_params(spp=1, first_sample=0, seed=1, max_depth=4))
    again = plugin.render_v4(v4.desc, params_v4(2, 0, seed=3, max_depth=8), v4.width, v4.height)
    np.testing.assert_array_equal(first, again)
    ref, _, _ = orc.OracleScene(v4.tabulated_scene(s)).render(
        pa.render_params(spp=2, first_sample=0, seed=3, max_depth=8, rr_start=1), variance=False)
    np.testing.assert_allclose(again.sum(), ref.sum(), rtol=5e-3)


def test_v4_render_adds_into_output(plugin, host):
    s = host.generate(1, 32, 32)
    v4 = V4Scene(host, s)
    a = plugin.render_v4(v4.desc, params_v4(2, 0, max_depth=8), v4.width, v4.height)
    b = plugin.render_v4(v4.desc, params_v4(2, 2, max_depth=8), v4.width, v4.height)
    both = plugin.render_v4(v4.desc, params_v4(2, 0, max_depth=8), v4.width, v4.height)
    plugin.render_v4(v4.desc, params_v4(2, 2, max_depth=8), v4.width, v4.height, out=both)
    np.testing.assert_allclose(both, a + b, rtol=1e-6, atol=1e-12)
    # sample streams are a function of the sample index: one 4-spp call equals two 2-spp calls
    four = plugin.render_v4(v4.desc, params_v4(4, 0, max_depth=8), v4.width, v4.height)
    np.testing.assert_allclose(four, a + b, rtol=1e-5, atol=1e-9)


def test_v4_render_errors(plugin, host):
    s = host.generate(1, 16, 16)
    v4 = V4Scene(host, s)
    v4.desc.apiVersion = 3
    with pytest.raises(RuntimeError, match="version"):
        plugin.render_v4(v4.desc, params_v4(1, 0), v4.width, v4.height)
    v4.desc.apiVersion = 4
    bad = s.materials.copy()
    bad["type"][0] = 9
    v4.desc.materials = bad.ctypes.data
    with pytest.raises(RuntimeError, match="material type"):
        plugin.render_v4(v4.desc, params_v4(1, 0), v4.width, v4.height)
    bad["type"][0] = pa.MAT_COATED_DIFFUSE  # a coat needs its IOR spectrum
    bad["coatEta"][0] = -1
    with pytest.raises(RuntimeError, match="coat spectra"):
        plugin.render_v4(v4.desc, params_v4(1, 0), v4.width, v4.height)
