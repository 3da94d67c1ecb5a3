"""Benchmark configuration 5 (round-brilliant diamonds with Sellmeier dispersion over rough
silver, a small 5000 K area light plus a constant environment, thin-lens camera) on the MI355X
plugin against the oracle, at reduced size: per-sample radiance from the same streams and the
converged image within 3 sigma.

Paths through a diamond take up to 32 specular events; a thin-lens origin or refraction that
differs in its last ulp (device vs host trigonometry) moves the exit point and ends in another
facet, so more samples diverge than on the diffuse scenes: the per-sample bar is 95% here."""
import pytest

import oracle as orc
import prender_amd as pa
from test_pt_gpu import check_samples, splat_correlation, three_sigma

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5(host):
    return host.generate(5, 128, 72, 0.25)


def test_sample_radiance_gems(plugin, c5):
    for si in (0, 9):
        check_samples(f"c5 si={si}", c5, pa.render_params(spp=1, first_sample=si, seed=1, max_depth=32), plugin, 0.95)


def test_image_3sigma_gems(plugin, c5):
    spp = 64
    g_sum, g_sq = plugin.render_ex(c5, pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=32))
    c_sum, c_sq, _ = orc.OracleScene(c5).render(pa.render_params(spp=spp, first_sample=spp, seed=2, max_depth=32))
    three_sigma("c5 image", g_sum, g_sq, spp, c_sum, c_sq, spp, splat_correlation(c5))
