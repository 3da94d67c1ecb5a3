"""Path-level pins of the oracle's path loop (its restatement of src/integrators/path.cpp:113-238),
from what the mathematics fixes, on the CPU.

The component goldens pin every function the loop calls; these pin the loop itself: how it
combines emission, next-event estimation and BSDF sampling under MIS, the spectral-MIS ratio
average, the eta scale and Russian roulette. Each would be off by more than its bar if one term were
dropped or mis-weighted:
- white furnaces (tests/pathscenes.py): mean XYZ = the D65 white point, Y within 1% (the reference's
  own bar, tests/test_render.cpp:85-88) and the chromaticity within 1%;
- an emissive Lambertian enclosure: Y = Le / (1 - rho) within 1%, and at depth cap d the truncated
  series Le (1 + rho + ... + rho^d) within 1%;
- a constant-density grid medium renders like the equivalent homogeneous medium within 3%
  (tests/test_media.cpp:96-115);
- the firefly clamp: every clamped pixel's Y equals the clamp within 1e-4 and keeps its colour
  (x / y, z / y) within 1e-4 (tests/test_quality.cpp:175-192).
"""
import numpy as np
import pytest

import oracle as orc
import pathscenes as ps
import prender_amd as pa

FURNACE_SPP = {"diffuse": 64, "glass": 64, "glass_bk7": 512, "medium": 256, "grid": 256}


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("pathscenes"))


@pytest.mark.parametrize("name", list(FURNACE_SPP))
def test_white_furnace(host, tmpdir, name):
    scene = ps.load(host, tmpdir, "furnace_" + name, ps.furnace_docs(tmpdir)[name])
    spp = FURNACE_SPP[name]
    s, sq, _ = orc.OracleScene(scene).render(pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=-1))
    mean, se = ps.mean_and_se(s, sq, spp)
    print(f"furnace {name}: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1.0) < 0.01, f"{name}: mean Y {mean[1]:.4f} is not 1 within 1%"
    assert abs(mean[1] - 1.0) < 5 * se[1] + 1e-3, f"{name}: mean Y {mean[1]:.4f} is {abs(mean[1] - 1) / se[1]:.1f} se off 1"
    for c in (0, 2):
        assert abs(mean[c] / mean[1] / ps.D65_XYZ[c] - 1) < 0.01, f"{name}: chromaticity off D65 in channel {c}"


def test_emissive_enclosure_depth_series(host, tmpdir):
    """With a depth cap d the path loop adds the emission of vertices 0..d (path.cpp:198-208: a
    vertex's emission is added before its depth check), so the enclosure sees the truncated Neumann
    series Le (1 + rho + ... + rho^d). Each term beyond the first is the MIS sum of a next-event
    estimate and a BSDF-sampled emission hit; losing or double-counting either half moves the mean by
    a whole term."""
    rho, le = 0.5, 0.5
    scene = ps.load(host, tmpdir, "enclosure_series", ps.enclosure_doc(rho, le))
    for d in (1, 2, 3):
        s, sq, _ = orc.OracleScene(scene).render(pa.render_params(spp=64, first_sample=0, seed=5, max_depth=d))
        mean, se = ps.mean_and_se(s, sq, 64)
        want = le * sum(rho ** k for k in range(d + 1))
        print(f"enclosure depth {d}: Y = {mean[1]:.5f} +- {se[1]:.5f} (want {want})", flush=True)
        assert abs(mean[1] / want - 1) < 0.01


def test_emissive_enclosure(host, tmpdir):
    for rho, le, spp in ((0.5, 0.5, 64), (0.8, 0.2, 256)):
        scene = ps.load(host, tmpdir, f"enclosure_{rho}", ps.enclosure_doc(rho, le))
        for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
            p = pa.render_params(spp=spp, first_sample=0, seed=3, max_depth=-1, light_sampler=sampler)
            s, sq, _ = orc.OracleScene(scene).render(p)
            mean, se = ps.mean_and_se(s, sq, spp)
            want = le / (1 - rho)
            print(f"enclosure rho={rho} Le={le} sampler={sampler}: Y = {mean[1]:.5f} +- {se[1]:.5f} (want {want})",
                  flush=True)
            assert abs(mean[1] / want - 1) < 0.01
            assert abs(mean[0] / mean[1] / ps.D65_XYZ[0] - 1) < 0.01


def test_constant_grid_matches_homogeneous(host, tmpdir):
    docs = ps.fog_docs(tmpdir)
    spp = 256
    y = {}
    for name, doc in docs.items():
        scene = ps.load(host, tmpdir, "fog_" + name, doc)
        s, sq, _ = orc.OracleScene(scene).render(pa.render_params(spp=spp, first_sample=0, seed=7, max_depth=-1))
        mean, se = ps.mean_and_se(s, sq, spp)
        y[name] = (mean[1], se[1])
    print(f"fog: homogeneous Y = {y['homogeneous'][0]:.5f} +- {y['homogeneous'][1]:.5f}, "
          f"grid Y = {y['grid'][0]:.5f} +- {y['grid'][1]:.5f}", flush=True)
    assert y["grid"][0] == pytest.approx(y["homogeneous"][0], rel=0.03)


def test_firefly_clamp_exact_and_keeps_colour(host, tmpdir):
    """A furnace whose environment is 1000 times brighter: every sample's Y is far above the clamp, so
    with one sample per pixel and a box filter each pixel holds exactly one clamped sample."""
    doc = ps.furnace_docs(tmpdir)["diffuse"]
    doc["lights"]["sky"]["intensity"] = 1000
    scene = ps.load(host, tmpdir, "bright_furnace", doc)
    o = orc.OracleScene(scene)
    raw, _, _ = o.render(pa.render_params(spp=1, first_sample=5, seed=1, max_depth=-1), variance=False)
    clamped, _, _ = o.render(pa.render_params(spp=1, first_sample=5, seed=1, max_depth=-1, clamp=2.0), variance=False)
    raw, clamped = raw.reshape(-1, 3), clamped.reshape(-1, 3)
    lit = raw[:, 1] > 0
    assert lit.mean() > 0.99
    assert raw[lit, 1].min() > 100
    np.testing.assert_allclose(clamped[lit, 1], 2.0, rtol=1e-4)
    for c in (0, 2):
        np.testing.assert_allclose(clamped[lit, c] / clamped[lit, 1], raw[lit, c] / raw[lit, 1], rtol=1e-4)
    # off by default: clamp 0 leaves the film untouched
    again, _, _ = o.render(pa.render_params(spp=1, first_sample=5, seed=1, max_depth=-1, clamp=0.0), variance=False)
    assert np.array_equal(again.reshape(-1, 3), raw)
