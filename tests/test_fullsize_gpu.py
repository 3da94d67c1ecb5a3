"""The headline workload at its full size: configuration 2 (fBm terrain + 200 icospheres, 2.02M
triangles, 64 blackbody quad lights, light-BVH next-event estimation) at 1920x1080, max depth 16,
Russian roulette from depth 1, the parameters the throughput benchmark renders with.

- Same-samples film parity: one sample per pixel from the MI355X and from the oracle with the same
  (pixel, sample index, seed) streams. The films differ only where a path diverged (a branch
  decided within float round-off, as in the small-film tests).
- Additivity over sample ranges: a 2-spp call equals the 1-spp calls at sample 0 and sample 1
  added together, a property that holds at any size and covers the batch, queue and film
  bookkeeping at full occupancy.
"""
import numpy as np
import pytest

import oracle as orc
import prender_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_full(host):
    return host.generate(2, 1920, 1080)


def _params(spp, first):
    return pa.render_params(spp=spp, first_sample=first, seed=1, max_depth=16, rr_start=1)


def test_full_size_film_matches_oracle_same_samples(plugin, c2_full):
    s = c2_full
    assert (s.width, s.height) == (1920, 1080) and len(s.prims) > 2_000_000
    g_sum, _ = plugin.render_ex(s, _params(1, 0))
    c_sum, _, secs = orc.OracleScene(s).render(_params(1, 0), threads=16)  # the host's CPU share
    assert np.all(np.isfinite(g_sum))
    rel = np.abs(g_sum - c_sum) / np.maximum(np.abs(c_sum), 1e-3)
    close = np.mean(rel <= 1e-3)
    print(f"c2 1920x1080 1 spp: pixels within 1e-3 = {close:.5f}, oracle {secs:.1f} s", flush=True)
    assert close >= 0.97
    np.testing.assert_allclose(g_sum.sum(), c_sum.sum(), rtol=5e-3)


def test_full_size_films_add_over_sample_ranges(plugin, c2_full):
    s = c2_full
    a, a_sq = plugin.render_ex(s, _params(1, 0))
    b, b_sq = plugin.render_ex(s, _params(1, 1))
    ab, ab_sq = plugin.render_ex(s, _params(2, 0))
    # The FP64 sums of the same samples, added in another order: equal to rounding.
    np.testing.assert_allclose(ab, a + b, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ab_sq, a_sq + b_sq, rtol=1e-9, atol=1e-12)
    assert not np.array_equal(a, b), "sample 1 must differ from sample 0"
