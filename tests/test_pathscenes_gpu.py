"""The closed-form scenes of tests/pathscenes.py rendered by the MI355X plugin.

The CPU suite (test_oracle_path.py) shows that the oracle's path loop (its restatement of
src/integrators/path.cpp:113-238) renders what the mathematics fixes. Here the HIP path is held to
the same two bars:
- per sample, the GPU's 4-wavelength radiance equals the oracle's for the same RNG stream
  (within 1e-4 relative) on every furnace, the emissive enclosure and the fog pair;
- per image, the GPU's own film has the closed-form mean: white furnaces at the D65 white point
  (Y within 1%, tests/test_render.cpp:85-88), the enclosure at Le / (1 - rho), the constant grid
  within 3% of the homogeneous medium (tests/test_media.cpp:96-115);
- the firefly clamp (src/gpu/prender_gpu.cpp:1550-1552) with a nonzero value: every clamped pixel's
  Y equals the clamp within 1e-4 and keeps its colour (tests/test_quality.cpp:175-192), and the
  clamped GPU film equals the clamped oracle film pixel for pixel.
"""
import numpy as np
import pytest

import oracle as orc
import pathscenes as ps
import prender_amd as pa
from test_pt_gpu import check_samples

pytestmark = pytest.mark.gpu

FURNACE_SPP = {"diffuse": 256, "glass": 256, "glass_bk7": 1024, "medium": 512, "grid": 512}
# Grid media hash the ray's float bytes for their tracking stream, so a last-ulp difference in a
# position changes every later decision of that path. The grid furnace's medium does not absorb and
# the depth is unlimited, so its paths hold tens of tracking events each; at the measured ~1% chance
# per event that a free-flight distance differs in its last ulp (the CPU's logf against the
# device's log), about 85% of its samples keep every event bit-identical. Its mean (Y = 1 within 1%)
# is the decisive check.
SAMPLE_BAR = {"diffuse": 0.995, "glass": 0.995, "glass_bk7": 0.99, "medium": 0.99, "grid": 0.80}


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("pathscenes_gpu"))


@pytest.mark.parametrize("name", list(FURNACE_SPP))
def test_white_furnace_gpu(plugin, host, tmpdir, name):
    scene = ps.load(host, tmpdir, "furnace_" + name, ps.furnace_docs(tmpdir)[name])
    check_samples(f"furnace {name}", scene, pa.render_params(spp=1, first_sample=3, seed=1, max_depth=-1),
                  plugin, SAMPLE_BAR[name])
    spp = FURNACE_SPP[name]
    s, sq = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=100, seed=2, max_depth=-1))
    mean, se = ps.mean_and_se(s, sq, spp)
    print(f"gpu furnace {name}: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1.0) < 0.01, f"{name}: mean Y {mean[1]:.4f} is not 1 within 1%"
    assert abs(mean[1] - 1.0) < 5 * se[1] + 1e-3
    for c in (0, 2):
        assert abs(mean[c] / mean[1] / ps.D65_XYZ[c] - 1) < 0.01, f"{name}: chromaticity off D65 in channel {c}"


def test_emissive_enclosure_gpu(plugin, host, tmpdir):
    rho, le = 0.8, 0.2
    scene = ps.load(host, tmpdir, "enclosure_gpu", ps.enclosure_doc(rho, le))
    for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
        check_samples(f"enclosure sampler={sampler}", scene,
                      pa.render_params(spp=1, first_sample=1, seed=3, max_depth=-1, light_sampler=sampler),
                      plugin, 0.99)
        spp = 512
        s, sq = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=64, seed=3, max_depth=-1,
                                                         light_sampler=sampler))
        mean, se = ps.mean_and_se(s, sq, spp)
        want = le / (1 - rho)
        print(f"gpu enclosure sampler={sampler}: Y = {mean[1]:.5f} +- {se[1]:.5f} (want {want})", flush=True)
        assert abs(mean[1] / want - 1) < 0.01


def test_enclosure_depth_series_gpu(plugin, host, tmpdir):
    rho, le = 0.5, 0.5
    scene = ps.load(host, tmpdir, "enclosure_series_gpu", ps.enclosure_doc(rho, le))
    for d in (1, 2, 3):
        s, sq = plugin.render_ex(scene, pa.render_params(spp=128, first_sample=0, seed=5, max_depth=d))
        mean, se = ps.mean_and_se(s, sq, 128)
        want = le * sum(rho ** k for k in range(d + 1))
        print(f"gpu enclosure depth {d}: Y = {mean[1]:.5f} +- {se[1]:.5f} (want {want})", flush=True)
        assert abs(mean[1] / want - 1) < 0.01


def test_constant_grid_matches_homogeneous_gpu(plugin, host, tmpdir):
    y = {}
    for name, doc in ps.fog_docs(tmpdir).items():
        scene = ps.load(host, tmpdir, "fog_gpu_" + name, doc)
        check_samples(f"fog {name}", scene, pa.render_params(spp=1, first_sample=2, seed=7, max_depth=-1),
                      plugin, SAMPLE_BAR["medium" if name == "homogeneous" else "grid"])
        spp = 1024
        s, sq = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=16, seed=7, max_depth=-1))
        mean, se = ps.mean_and_se(s, sq, spp)
        y[name] = (mean[1], se[1])
    print(f"gpu fog: homogeneous Y = {y['homogeneous'][0]:.5f} +- {y['homogeneous'][1]:.5f}, "
          f"grid Y = {y['grid'][0]:.5f} +- {y['grid'][1]:.5f}", flush=True)
    assert y["grid"][0] == pytest.approx(y["homogeneous"][0], rel=0.03)


def test_firefly_clamp_gpu(plugin, host, tmpdir):
    """A furnace under an environment 1000 times brighter: one sample per pixel, box filter, so every
    pixel holds exactly one sample and every lit sample's Y is far above the clamp."""
    doc = ps.furnace_docs(tmpdir)["diffuse"]
    doc["lights"]["sky"]["intensity"] = 1000
    scene = ps.load(host, tmpdir, "bright_furnace_gpu", doc)
    o = orc.OracleScene(scene)
    for clamp in (2.0, 37.5):
        p = pa.render_params(spp=1, first_sample=5, seed=1, max_depth=-1, clamp=clamp)
        raw, _ = plugin.render_ex(scene, pa.render_params(spp=1, first_sample=5, seed=1, max_depth=-1))
        g, _ = plugin.render_ex(scene, p)
        c, _, _ = o.render(p, variance=False)
        raw, g, c = raw.reshape(-1, 3), g.reshape(-1, 3), c.reshape(-1, 3)
        lit = raw[:, 1] > 0
        assert lit.mean() > 0.99 and raw[lit, 1].min() > 100 * clamp / 37.5
        np.testing.assert_allclose(g[lit, 1], clamp, rtol=1e-4)
        for ch in (0, 2):
            np.testing.assert_allclose(g[lit, ch] / g[lit, 1], raw[lit, ch] / raw[lit, 1], rtol=1e-4)
        # the clamped films of the two sides, pixel for pixel
        close = np.mean(np.abs(g - c) <= 1e-4 * np.maximum(np.abs(c), 1e-6))
        print(f"gpu clamp {clamp}: pixels within 1e-4 of the oracle's clamped film = {close:.4f}", flush=True)
        assert close >= 0.99
