"""Run-time tuning on the MI355X: every launch shape and queue capacity renders the same film.

The path of every camera sample is a pure function of (pixel, sample index, seed), so the grids of
the traversal and shading kernels and replaying captured bounce graphs instead of launching eagerly
must give the film of the default settings bit for bit, and so must sorting the material queues by
material and medium (the order the material stage takes its paths in) and sorting the next
bounce's rays by direction and origin (the order they are traced in), and a smaller queue capacity (paths in
flight per batch, which splits a call into batches) the same film up to the grouping of its FP64
sums. Captured graphs are replayed many times per call
(one replay per graph_rounds bounces), with the eager renders of the same sample ranges between
them, and every call is checked.
"""
import os

import numpy as np
import pytest

import prender_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture()
def defaults(plugin):
    saved = {k: plugin.tune_get(k) for k in plugin.TUNE_KEYS}
    yield saved
    for k, v in saved.items():
        plugin.tune_set(k, v)


def _film(plugin, scene, first, spp=4):
    s, _ = plugin.render_ex(scene, pa.render_params(spp=spp, first_sample=first, seed=1, max_depth=8, rr_start=3),
                            variance=False)
    return s


def test_graph_replays_equal_eager_launches(plugin, host, defaults):
    s = host.generate(1, 96, 96)
    for rounds in (2, 4):
        for call in range(3):
            plugin.tune_set("graph_rounds", 0)
            eager = _film(plugin, s, 8 * call)
            plugin.tune_set("graph_rounds", rounds)
            graphed = _film(plugin, s, 8 * call)
            assert np.array_equal(graphed, eager), f"graph_rounds={rounds}, call {call}: films differ"
        print(f"graph_rounds={rounds}: 3 calls equal the eager films", flush=True)


@pytest.mark.parametrize("key,value", [("trace_blocks_per_cu", 3), ("shade_blocks_per_cu", 5),
                                       ("path_budget", 65536), ("material_sort", 1), ("ray_sort", 1)])
def test_launch_shape_and_queue_capacity_keep_the_film(plugin, host, defaults, key, value):
    s = host.generate(2, 160, 90)
    ref = _film(plugin, s, 0, spp=16)
    plugin.tune_set(key, value)
    assert plugin.tune_get(key) == value
    film = _film(plugin, s, 0, spp=16)
    if key == "path_budget":
        # 16 passes in batches of 4: the same samples, their FP64 film sums grouped per batch
        np.testing.assert_allclose(film, ref, rtol=1e-12, atol=1e-300)
    else:
        assert np.array_equal(film, ref)


@pytest.mark.parametrize("key", ["material_sort", "ray_sort"])
def test_sorts_keep_the_film_with_media_coats_and_guiding(plugin, host, defaults, key):
    """The material sort and the ray sort on scenes where every queue is in use: the sample
    scene's glass, metals, coats, sheen and subsurface media (the phase-function queue sorted by
    medium; interface crossings re-enter the path queue from the surface stage), and the room with
    its learned guide (the guided diffuse queue is inferred and shaded in sorted order)."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample_scene.npz"))
    for name in ("studio", "studio_fog"):
        scene = pa.flat_scene(g, name)
        prm = pa.render_params(spp=2, first_sample=5, seed=1, max_depth=scene.defaults["maxDepth"], rr_start=1)
        films = []
        for srt in (0, 1):
            plugin.tune_set(key, srt)
            films.append(plugin.render_ex(scene, prm, variance=False)[0])
        assert films[0].any() and np.array_equal(films[0], films[1]), f"{name}: {key} changed the film"
    room = host.generate(4, 96, 54, 0.25)
    films = []
    for srt in (0, 1):
        plugin.tune_set(key, srt)
        prm = pa.render_params(spp=4, first_sample=0, seed=1, max_depth=8, flags=pa.RENDER_GUIDE | pa.RENDER_GUIDE_RESET)
        films.append(plugin.render_ex(room, prm, variance=False)[0])
    assert films[0].any() and np.array_equal(films[0], films[1]), f"guided room: {key} changed the film"
    print(f"{key}: sample scene (studio, fog) and guided room films equal the unsorted ones", flush=True)


def test_out_of_range_values_are_refused(plugin, defaults):
    for key, value in (("trace_blocks_per_cu", 0), ("graph_rounds", 3), ("guide_blend", 1.5), ("no_such_key", 1)):
        with pytest.raises(ValueError):
            plugin.tune_set(key, value)
