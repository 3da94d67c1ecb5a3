"""Image environment lights and the Preetham sun-sky rendered by the MI355X plugin.

The CPU suite (test_envmap.py) pins the oracle's image environment (src/lights/light.cpp:215-297)
against the host's Distribution2D tables and the closed forms of a constant map. Here the HIP path
is held to:
- per sample, the GPU's radiance equals the oracle's for the same RNG stream on a scene lit only by
  a studio map (a bright sun blob, a rotated map, diffuse, conductor and glass objects) and on the
  sun-sky scene, under both light samplers;
- per image, the GPU's own film of a reflectance-1 sphere under a constant white map is the white
  furnace: Y = 1 within 1% at the D65 white point.
"""
import numpy as np
import pytest

import envscenes as es
import pathscenes as ps
import prender_amd as pa
from test_pt_gpu import check_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("envmap_gpu"))


@pytest.mark.parametrize("sampler", [pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER])
def test_studio_map_samples(plugin, host, tmpdir, sampler):
    es.write_pfm(f"{tmpdir}/studio.pfm", es.studio_map(64, 32))
    scene = es.load(host, tmpdir, "studio", es.env_doc(f"{tmpdir}/studio.pfm", rotation=35, film=(48, 48)))
    for si in (0, 5):
        check_samples(f"studio map sampler={sampler} si={si}", scene,
                      pa.render_params(spp=1, first_sample=si, seed=4, max_depth=8, light_sampler=sampler),
                      plugin, 0.99)


@pytest.mark.parametrize("sampler", [pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER])
def test_sun_sky_samples(plugin, host, tmpdir, sampler):
    scene = es.load(host, tmpdir, "sunsky", es.sun_sky_doc(resolution=128))
    check_samples(f"sun-sky sampler={sampler}", scene,
                  pa.render_params(spp=1, first_sample=2, seed=6, max_depth=8, light_sampler=sampler), plugin, 0.99)


def test_white_map_furnace_gpu(plugin, host, tmpdir):
    es.write_pfm(f"{tmpdir}/white.pfm", es.constant_map((1, 1, 1)))
    scene = es.load(host, tmpdir, "furnace", es.furnace_env_doc(f"{tmpdir}/white.pfm"))
    check_samples("image-map furnace", scene, pa.render_params(spp=1, first_sample=0, seed=1, max_depth=-1),
                  plugin, 0.995)
    spp = 512
    s, sq = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=64, seed=2, max_depth=-1))
    mean, se = ps.mean_and_se(s, sq, spp)
    print(f"gpu image-map furnace: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1) < 0.01
    for c in (0, 2):
        assert abs(mean[c] / mean[1] / ps.D65_XYZ[c] - 1) < 0.01
