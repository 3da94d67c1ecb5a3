"""Participating media on the MI355X plugin against the oracle: the configuration-3 scene (a
heterogeneous grid medium behind an interface box, a subsurface sphere with a homogeneous
interior behind a smooth IOR 1.33 boundary, one distant light) at reduced grid resolution.

Grid media use spectral tracking for distance sampling and ratio tracking for transmittance
(reference src/media/medium.cpp:205-285), driven by a generator hashed from the ray's float bytes,
so a sample agrees only while its whole path is bit-identical; the per-sample agreement bound is
looser than for surfaces, and the converged-image test carries the statistical claim.
"""
import numpy as np
import pytest

import oracle as orc
import prender_amd as pa
from test_pt_gpu import check_samples, three_sigma

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(host):
    return host.generate(3, 96, 54, 0.125)


def test_sample_radiance_media(plugin, c3):
    for si in (0, 5):
        check_samples(f"c3 si={si}", c3, pa.render_params(spp=1, first_sample=si, seed=1, max_depth=256), plugin, 0.95)


def test_image_3sigma_media(plugin, c3):
    spp = 128
    g_sum, g_sq = plugin.render_ex(c3, pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=256))
    c_sum, c_sq, _ = orc.OracleScene(c3).render(pa.render_params(spp=spp, first_sample=spp, seed=2, max_depth=256))
    three_sigma("c3 image", g_sum, g_sq, spp, c_sum, c_sq, spp)


def test_media_power_light_sampler_and_depth_cap(plugin, c3):
    p = pa.render_params(spp=1, first_sample=2, seed=3, max_depth=4, light_sampler=pa.LIGHT_SAMPLER_POWER)
    check_samples("c3 depth 4 power", c3, p, plugin, 0.95)
