"""A scene loaded from a .prscene.json file, rendered by the MI355X plugin against the oracle.

The studio fixture (tests/scenes/studio) holds every closure family (Lambert with a checker,
conductors by metal and by reflectance, rough and dispersive dielectrics, a subsurface sphere,
coated diffuse with sheen, coated conductor), an OBJ and a PLY mesh, rect / sphere / triangle area
lights with lobes and power-given emission, a distant light and a constant sky, seen through a thin
lens with a Gaussian filter; the haze rig adds a homogeneous medium behind an interface box that
holds the camera. The loader runs on the host (libprender_host.so); the per-sample check compares
each camera sample's radiance with the oracle's for the same (pixel, sample index, seed). The bar
is that of the coated materials, whose random walk agrees with the CPU only while every direction
it produces is bit-identical.
"""
import os

import pytest

import prender_amd as pa
from test_pt_gpu import check_samples

pytestmark = pytest.mark.gpu

STUDIO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes", "studio", "scene.prscene.json")
SAMPLE_BAR = 0.97


def test_sample_radiance_studio_file(plugin, host):
    s = host.load_scene(STUDIO)
    print("studio warnings:", s.warnings, flush=True)
    for si in (0, 5):
        check_samples(f"studio si={si}", s, pa.render_params(spp=1, first_sample=si, seed=s.defaults["seed"],
                                                             max_depth=s.defaults["maxDepth"], rr_start=1,
                                                             light_sampler=pa.LIGHT_SAMPLER_POWER),
                      plugin, SAMPLE_BAR)


def test_sample_radiance_studio_file_in_haze(plugin, host):
    s = host.load_scene(STUDIO, rigs=["lamps", "haze"])
    assert s.desc.cameraMedium >= 0
    check_samples("studio+haze", s, pa.render_params(spp=1, first_sample=2, seed=3, max_depth=24, rr_start=1),
                  plugin, SAMPLE_BAR)


def test_sample_radiance_spot_lights(plugin, host):
    """Spot lights (smoothstep falloff, cone bounds in the light BVH) beside point and area lights."""
    s = host.load_scene(os.path.join(os.path.dirname(STUDIO), "..", "spots", "scene.prscene.json"))
    assert (s.lights["kind"] == pa.LIGHT_SPOT).sum() == 3
    for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
        check_samples(f"spots sampler={sampler}", s, pa.render_params(spp=1, first_sample=1, seed=5, max_depth=6,
                                                                     rr_start=1, light_sampler=sampler),
                      plugin, 0.99)


def test_sample_radiance_goniometric_lights(plugin, host):
    """IES (goniometric) lights of the three symmetries, with a spot and a rect light."""
    s = host.load_scene(os.path.join(os.path.dirname(STUDIO), "..", "ies", "scene.prscene.json"))
    assert (s.lights["kind"] == pa.LIGHT_GONIO).sum() == 3
    for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
        check_samples(f"ies sampler={sampler}", s, pa.render_params(spp=1, first_sample=2, seed=9, max_depth=6,
                                                                   rr_start=1, light_sampler=sampler),
                      plugin, 0.99)


def test_sample_radiance_sun(plugin, host):
    """The sun: cone sampling over its disk, emission met by escaping paths inside the disk with MIS,
    beside a constant sky and a rect light, through a glass sphere."""
    s = host.load_scene(os.path.join(os.path.dirname(STUDIO), "..", "sun", "scene.prscene.json"))
    assert (s.lights["kind"] == pa.LIGHT_SUN).sum() == 1
    for sampler in (pa.LIGHT_SAMPLER_BVH, pa.LIGHT_SAMPLER_POWER):
        check_samples(f"sun sampler={sampler}", s, pa.render_params(spp=1, first_sample=3, seed=2, max_depth=8,
                                                                   rr_start=1, light_sampler=sampler),
                      plugin, 0.99)


def test_sample_radiance_reference_sample_scene(plugin):
    """The reference's sample scene (scenes/sphere-pyramid as the loader flattens it: glass with
    dispersion, metals, subsurface, coats, sheen, an emissive sphere; studio rig, and studio + fog),
    at a quarter of its film. The per-sample bar is that of the coated materials."""
    import numpy as np
    g = np.load(os.path.join(os.path.dirname(STUDIO), "..", "..", "golden", "sample_scene.npz"))
    for name in ("studio", "studio_fog"):
        full = pa.flat_scene(g, name)
        cam = pa.CameraEx.from_buffer_copy(bytes(full.camera))
        cam.width, cam.height = 240, 135
        s = pa.Scene(full.prims, full.materials, full.spectra, full.knots, full.lights, full.media, full.grid, cam,
                     full.desc.cameraMedium, full.defaults)
        check_samples(f"sample scene {name}", s, pa.render_params(spp=1, first_sample=4, seed=1, max_depth=64,
                                                                 rr_start=1), plugin, SAMPLE_BAR)
