"""Image textures, normal maps and bump maps rendered by the MI355X plugin.

The CPU suite pins the oracle's texture lookup and shading perturbation bit for bit against the
reference's compiled code (test_oracle_golden.py) and its path loop with them against closed forms
(test_textures.py). Here the HIP path is held to:
- per sample, the GPU's radiance equals the oracle's for the same RNG stream on a set with an image
  albedo tiled with mirror wrap, a normal-mapped sphere (flipped green, strength 1.5) and a
  bump- and normal-mapped rough conductor with an image reflectance and a sheened coated-diffuse
  silk with a PNG reflectance and a PNG bump weave, under an image environment and a quad light,
  with both light samplers;
- per image, the closed forms on the GPU's own films: a constant normal map scales an overhead
  distant light's direct radiance by t_z / |t|, a linear bump ramp by L_u / sqrt(s^2 + L_u^2), and
  a white image texture on a reflectance-1 sphere under a white map is a white furnace.
"""
import numpy as np
import pytest

import envscenes as es
import pathscenes as ps
import prender_amd as pa
import texscenes as ts
from test_pt_gpu import check_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("textures_gpu"))


@pytest.mark.parametrize("sampler", [pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER])
def test_textured_set_samples(plugin, host, tmpdir, sampler):
    scene = es.load(host, tmpdir, "textured", ts.textured_objects_doc(tmpdir))
    for si in (0, 3):
        check_samples(f"textured set sampler={sampler} si={si}", scene,
                      pa.render_params(spp=1, first_sample=si, seed=8, max_depth=8, light_sampler=sampler), plugin,
                      0.99)


def _film(plugin, scene, spp=4):
    s, _ = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=0, seed=3, max_depth=1))
    return s


def test_normal_and_bump_maps_closed_forms_gpu(plugin, host, tmpdir):
    c = np.array([0.8, 0.5, 0.7], np.float32)
    t = 2 * c - 1
    plain = _film(plugin, es.load(host, tmpdir, "np0", ts.plane_doc(tmpdir)))
    mapped = _film(plugin, es.load(host, tmpdir, "np1", ts.plane_doc(tmpdir, normal=es.constant_map(c, 3, 3))))
    r = mapped / plain
    print(f"gpu normal map: ratio {r.min():.6f}..{r.max():.6f}, closed form {t[2] / np.linalg.norm(t):.6f}", flush=True)
    np.testing.assert_allclose(r, t[2] / np.linalg.norm(t), rtol=2e-5)
    size, scale = 16.0, 4.0
    plain = _film(plugin, es.load(host, tmpdir, "bp0", ts.plane_doc(tmpdir, size=size)))
    bumped = _film(plugin, es.load(host, tmpdir, "bp1", ts.plane_doc(tmpdir, size=size, bump=(ts.ramp(2), scale))))
    r = bumped / plain
    want = size / np.hypot(2 * scale, size)
    print(f"gpu bump ramp: ratio {r.min():.6f}..{r.max():.6f}, closed form {want:.6f}", flush=True)
    np.testing.assert_allclose(r, want, rtol=2e-3)


def test_white_texture_furnace_gpu(plugin, host, tmpdir):
    es.write_pfm(f"{tmpdir}/white_env.pfm", es.constant_map((1, 1, 1)))
    es.write_pfm(f"{tmpdir}/tex_white.pfm", es.constant_map((1, 1, 1), 4, 4))
    doc = es.furnace_env_doc(f"{tmpdir}/white_env.pfm")
    doc["textures"] = {"w": {"type": "image", "file": f"{tmpdir}/tex_white.pfm"}}
    doc["materials"]["m"]["reflectance"] = {"texture": "w"}
    scene = es.load(host, tmpdir, "tex_furnace", doc)
    check_samples("white-texture furnace", scene, pa.render_params(spp=1, first_sample=0, seed=1, max_depth=-1),
                  plugin, 0.995)
    spp = 512
    s, sq = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=64, seed=2, max_depth=-1))
    mean, se = ps.mean_and_se(s, sq, spp)
    print(f"gpu white-texture furnace: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1) < 0.01
    for k in (0, 2):
        assert abs(mean[k] / mean[1] / ps.D65_XYZ[k] - 1) < 0.01
