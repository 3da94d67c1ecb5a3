"""Scene files on the CPU: the .prscene.json loader (reference src/scene/scene_loader.cpp), its OBJ and
PLY readers (src/geometry/mesh_io.cpp) and the flattening into the renderer's scene that the reference's
GPU exporter performs (src/integrators/gpu_path.cpp:104-456).

The fixture tests/scenes/studio is this project's own scene in the reference's format. Its expected
values come from the format's definitions, written out here: transforms compose as translate x rotate x
scale; roughness maps to alpha = roughness^2; emission given as power becomes
power / (Y area lobe) with lobe = 2 pi / (n + 2) (x2 when two-sided); an f-stop gives the lens radius
(focal length / 1000) / (2 N); a rect light's emitting side faces its target. The reference's own
scene files are loaded too when the reference tree is present (they are not copied here).
"""
import json
import math
import os
import re
import struct

import numpy as np
import pytest

import prender_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
STUDIO = os.path.join(HERE, "scenes", "studio", "scene.prscene.json")
REF_SCENES = "/root/reference/scenes"


@pytest.fixture(scope="module")
def studio(host):
    return host.load_scene(STUDIO)


def tri_normal(p):
    a, b, c = (np.array(p[k], np.float64) for k in ("p0", "p1", "p2"))
    n = np.cross(b - a, c - a)
    return n, 0.5 * np.linalg.norm(n)


def spectrum_y(host, scene, idx):
    """Luminance Y of a scene spectrum: 1 nm quadrature of the host's evaluation of it against an
    independent CIE y-bar fit, normalised so that a constant 1 has Y = 1."""
    lam = np.arange(360, 831, dtype=np.float64)
    v = np.array([host.spectrum_value(scene.spectra[idx], scene.knots, float(l)) for l in lam])
    ybar = cie_y(lam)
    return float((v * ybar).sum() / ybar.sum())


def cie_y(lam):
    """CIE 1931 y-bar, Wyman, Sloan and Shirley's multi-lobe fit (JCGT 2013)."""
    def g(x, mu, s1, s2):
        return np.exp(-0.5 * ((x - mu) / np.where(x < mu, s1, s2)) ** 2)
    return 0.821 * g(lam, 568.8, 46.9, 40.5) + 0.286 * g(lam, 530.9, 16.3, 31.1)


def key(p):
    return tuple(round(float(x), 5) for x in p["p0"])


def sphere_prims(s):
    return s.prims[s.prims["type"] == pa.PRIM_SPHERE]


# ---- composition, counts, settings -----------------------------------------------------------

def test_studio_counts_rigs_and_settings(host, studio):
    s = studio
    # floor 2, five spheres, box 12, disk 12, octahedron 8 + 1 from the quad's fan, wedge 6, triangle 1;
    # the lamps rig adds the key rect (2) and the bulb sphere
    assert len(s.prims) == 2 + 5 + 12 + 12 + 9 + 6 + 1 + 2 + 1
    kinds = sorted(s.lights["kind"].tolist())
    # sign triangle, key rect (2), bulb, sun (distant), sky, spot; "off" is disabled
    assert kinds == sorted([pa.LIGHT_AREA] * 4 + [pa.LIGHT_DISTANT, pa.LIGHT_ENV, pa.LIGHT_SPOT])
    assert any("roughness textures" in w for w in s.warnings)
    assert not any("spot" in w for w in s.warnings)
    assert (s.width, s.height) == (96, 64)
    assert s.camera.filterType == pa.FILTER_GAUSSIAN and s.camera.filterRadius == pytest.approx(1.25)
    assert s.defaults == {"spp": 32, "maxDepth": 12, "rrStartDepth": 1, "seed": 11}
    assert s.settings == {"clamp": 40.0, "time_limit": 90.0, "russian_roulette": True, "light_bvh": False}
    # every area light's primitive points back at its light
    area = np.flatnonzero(s.lights["kind"] == pa.LIGHT_AREA)
    assert np.array_equal(s.prims["light"][s.lights["primitive"][area]], area)
    # rigs replace the active ones; an inline rig and no rig at all
    assert len(host.load_scene(STUDIO, rigs=[]).lights) == 1
    inline = host.load_scene(STUDIO, rigs=["inline"])
    assert sorted(inline.lights["kind"].tolist()) == [pa.LIGHT_AREA, pa.LIGHT_POINT]
    assert (host.load_scene(STUDIO, width=40, height=30).width, host.load_scene(STUDIO, width=40, height=30).height) == (40, 30)


def test_transforms_and_spheres(studio):
    sp = sphere_prims(studio)
    centres = {key(p): p["r"] for p in sp}
    assert centres[(-1.2, 0.35, 0.0)] == pytest.approx(0.35)
    assert centres[(0.4, 0.35, 0.3)] == pytest.approx(0.42)       # uniform scale 1.2
    assert centres[(0.0, 0.2, 1.1)] == pytest.approx(0.175)       # scale [0.5, 0.5, 0.5]
    assert centres[(1.8, 1.6, -1.0)] == pytest.approx(0.12)       # the bulb light
    assert len(centres) == 5 + 1  # the invisible object is not loaded


def test_box_disk_and_rect_areas(host, tmp_path):
    doc = {"geometry": {"b": {"type": "box", "size": [0.5, 0.4, 0.6]}, "d": {"type": "disk", "radius": 0.3, "segments": 12},
                        "r": {"type": "rect", "size": [8, 6]}},
           "materials": {"m": {"type": "diffuse"}},
           "objects": [{"geometry": "b", "material": "m", "transform": {"rotate": [[30, 0, 1, 0], [5, 1, 0, 0]]}},
                       {"geometry": "d", "material": "m"}, {"geometry": "r", "material": "m"}],
           "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
    f = tmp_path / "shapes.prscene.json"
    f.write_text(json.dumps(doc))
    s = host.load_scene(f)
    areas = np.array([tri_normal(p)[1] for p in s.prims])
    assert np.isclose(areas[:12].sum(), 2 * (0.5 * 0.4 + 0.4 * 0.6 + 0.5 * 0.6), rtol=1e-5)
    assert np.isclose(areas[12:24].sum(), 12 * 0.5 * 0.3 * 0.3 * math.sin(2 * math.pi / 12), rtol=1e-5)
    assert np.isclose(areas[24:].sum(), 48, rtol=1e-6)
    # box faces are wound outwards (the box is closed): normals point away from the centre
    for p in s.prims[:12]:
        n, _ = tri_normal(p)
        c = (np.array(p["p0"]) + np.array(p["p1"]) + np.array(p["p2"])) / 3
        assert np.dot(n, c) > 0
    # disk and rect face +y; both carry shading normals
    for p in s.prims[12:]:
        n, _ = tri_normal(p)
        assert n[1] > 0 and abs(n[0]) < 1e-6 and abs(n[2]) < 1e-6
        assert np.allclose(p["n0"], [0, 1, 0])
    assert "camera" not in " ".join(s.warnings)


# ---- materials and media ---------------------------------------------------------------------

def material_of(s, name_prim_index):
    return s.materials[s.prims["material"][name_prim_index]]


def test_material_parameters(host, studio):
    s = studio
    sp = sphere_prims(s)
    by_centre = {key(p): p for p in sp}
    gold = s.materials[by_centre[(-1.2, 0.35, 0.0)]["material"]]
    assert gold["type"] == pa.MAT_CONDUCTOR and gold["alphaX"] == pytest.approx(0.04) and gold["k"] >= 0
    glass = s.materials[by_centre[(-0.4, 0.35, 0.3)]["material"]]
    assert glass["type"] == pa.MAT_DIELECTRIC and glass["alphaX"] == pytest.approx(0.09) and glass["dispersive"] == 1
    assert s.spectra[glass["eta"]]["kind"] == pa.SPEC_SELLMEIER
    jade_p = by_centre[(0.4, 0.35, 0.3)]
    jade = s.materials[jade_p["material"]]
    assert jade["type"] == pa.MAT_DIELECTRIC and jade["alphaX"] == pytest.approx(0.01) and jade["dispersive"] == 0
    assert host.spectrum_value(s.spectra[jade["eta"]], s.knots, 550.0) == pytest.approx(1.6)
    velvet = s.materials[by_centre[(1.2, 0.35, 0.0)]["material"]]
    assert velvet["type"] == pa.MAT_COATED_DIFFUSE
    assert velvet["coatAlpha"] == pytest.approx(0.04) and velvet["thickness"] == pytest.approx(0.02)
    assert velvet["coatG"] == pytest.approx(0.3) and velvet["coatAlbedo"] >= 0
    assert velvet["sheenWeight"] == pytest.approx(0.8) and velvet["sheenRoughness"] == pytest.approx(0.35)
    water = s.materials[by_centre[(0.0, 0.2, 1.1)]["material"]]
    assert s.spectra[water["eta"]]["kind"] == pa.SPEC_CAUCHY and water["dispersive"] == 1
    eta = s.spectra[water["eta"]]
    assert host.spectrum_value(eta, s.knots, 500.0) == pytest.approx(1.3240 + 3089 / 500 ** 2, rel=1e-6)
    # the floor's checker keeps both colours and the u scale
    floor = s.materials[s.prims["material"][0]]
    assert floor["type"] == pa.MAT_DIFFUSE and floor["albedoB"] >= 0 and floor["checkerScale"] == pytest.approx(12)
    # artist conductor (reflectance) with anisotropic roughness
    mats = s.materials
    cu = mats[(mats["type"] == pa.MAT_CONDUCTOR) & (mats["k"] < 0)]
    assert len(cu) == 1 and cu[0]["alphaX"] == pytest.approx(0.09) and cu[0]["alphaY"] == pytest.approx(0.01)
    # coated conductor whose roughness is a texture: approximated by the default (smooth)
    cc = mats[mats["type"] == pa.MAT_COATED_CONDUCTOR]
    assert len(cc) == 1 and cc[0]["alphaX"] == 0 and cc[0]["coatAlpha"] == pytest.approx(0.0025)


def test_subsurface_medium_from_albedo_and_mfp(host, studio):
    s = studio
    jade_p = [p for p in sphere_prims(s) if np.allclose(p["p0"], [0.4, 0.35, 0.3])][0]
    m = s.media[jade_p["mediumInside"]]
    assert m["kind"] == pa.MEDIUM_HOMOGENEOUS and jade_p["mediumOutside"] == -1
    # sigma_t = 1 / mfp in every channel and sigma_s / sigma_t = the single-scattering albedo that
    # van de Hulst's relation gives for the multiple-scattering albedo; checked through an RGB-exact
    # channel: the unbounded uplift of a grey triple is a constant, so probe the green channel ratio
    alb, mfp = np.array([0.55, 0.9, 0.6]), np.array([0.05, 0.1, 0.06])
    a = np.clip(alb, 0, 0.999)
    alpha = np.clip(1 - (4.09712 + 4.20863 * a - np.sqrt(9.59217 + 41.6808 * a + 17.7126 * a * a)) ** 2, 0, 0.9999)
    # alpha is monotone in the albedo and lies in (0, 1) for albedos in (0, 1)
    assert np.all(np.diff(alpha[np.argsort(a)]) > 0) and np.all((alpha > 0) & (alpha < 1))
    sa = np.array([host.spectrum_value(s.spectra[m["sigmaA"]], s.knots, l) for l in (450.0, 550.0, 610.0)])
    ss = np.array([host.spectrum_value(s.spectra[m["sigmaS"]], s.knots, l) for l in (450.0, 550.0, 610.0)])
    st = sa + ss
    # extinction sits between the channels' 1/mfp, scattering dominates where the albedo is high
    assert np.all(st > 1 / mfp.max() * 0.9) and np.all(st < 1 / mfp.min() * 1.1)
    assert ss[1] / st[1] > ss[0] / st[0]


def test_media_resolution_with_haze_rig(host):
    s = host.load_scene(STUDIO, rigs=["lamps", "haze"])
    haze = s.desc.cameraMedium
    assert haze >= 0 and s.media[haze]["g"] == pytest.approx(0.5)
    is_box = s.materials["type"][s.prims["material"]] == pa.MAT_INTERFACE
    box = s.prims[is_box]  # the haze volume: the only interface object
    assert len(box) == 12 and np.all(box["mediumInside"] == haze) and np.all(box["mediumOutside"] == -1)
    rest = s.prims[~is_box]
    assert np.all(rest["mediumOutside"] == haze)
    water = [p for p in sphere_prims(s) if np.allclose(p["p0"], [0.0, 0.2, 1.1])][0]
    assert water["mediumInside"] not in (-1, haze) and water["mediumOutside"] == haze
    # without the haze nothing is outside a medium and the camera is in vacuum
    plain = host.load_scene(STUDIO)
    assert plain.desc.cameraMedium == -1 and np.all(plain.prims["mediumOutside"] == -1)


# ---- lights and emission -------------------------------------------------------------------------

def test_rect_light_faces_its_target_and_keeps_its_lobe(studio):
    s = studio
    key = s.lights[(s.lights["kind"] == pa.LIGHT_AREA) & (s.lights["cosPower"] == 3)]
    assert len(key) == 2 and np.all(key["scale"] == np.float32(14))
    pos, target = np.array([-2.5, 3.5, 2.0]), np.array([0, 0.3, 0])
    d = (target - pos) / np.linalg.norm(target - pos)
    total = 0
    for li in key:
        n, area = tri_normal(s.prims[li["primitive"]])
        assert np.allclose(n / np.linalg.norm(n), d, atol=1e-5)
        total += area
    assert total == pytest.approx(1.2 * 0.8, rel=1e-5)
    assert s.materials[s.prims[key[0]["primitive"]]["material"]]["type"] == pa.MAT_DIFFUSE


def test_power_emission_distant_and_environment(host, studio):
    s = studio
    sign = s.lights[(s.lights["kind"] == pa.LIGHT_AREA) & (s.lights["twoSided"] == 1)]
    assert len(sign) == 1
    area = tri_normal(s.prims[sign[0]["primitive"]])[1]
    assert area == pytest.approx(0.08, rel=1e-6)
    y = spectrum_y(host, s, sign[0]["spectrum"])
    assert y == pytest.approx(1.0, rel=2e-3)  # blackbodies are normalised to unit luminance
    lobe = 2 * math.pi / (0 + 2) * 2
    assert sign[0]["scale"] == pytest.approx(15 / (y * area * lobe), rel=3e-3)
    sun = s.lights[s.lights["kind"] == pa.LIGHT_DISTANT][0]
    d = -np.array([0.3, -1.0, -0.4])
    assert np.allclose(sun["direction"], d / np.linalg.norm(d), atol=1e-6) and sun["scale"] == pytest.approx(1.5)
    env = s.lights[s.lights["kind"] == pa.LIGHT_ENV][0]
    assert env["scale"] == pytest.approx(0.15)


def test_illuminant_rgb_white_is_d65_at_unit_luminance(host, tmp_path):
    """An illuminant RGB (1, 1, 1) uplifts to 2 * S(0.5) * D65 = D65 exactly (the sigmoid of a zero
    polynomial is 1/2), and D65 is normalised to Y = 1: its luminance must come out as 1, and a
    named "D65" spectrum must agree with it at every 5 nm node."""
    doc = {"lights": {"a": {"type": "environment", "color": [1, 1, 1]}, "b": {"type": "point", "position": [0, 1, 0],
                                                                             "spectrum": "D65", "power": 2}},
           "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
    f = tmp_path / "white.prscene.json"
    f.write_text(json.dumps(doc))
    s = host.load_scene(f)
    env = s.lights[s.lights["kind"] == pa.LIGHT_ENV][0]
    pt = s.lights[s.lights["kind"] == pa.LIGHT_POINT][0]
    assert spectrum_y(host, s, env["spectrum"]) == pytest.approx(1.0, rel=2e-3)
    # (at 830 nm the 5 nm table's lookup clamps its index just below the last node and blends
    # 1/1000 of the 825 nm value in; every node before it is exact)
    for lam in np.arange(360.0, 826.0, 5.0):
        a = host.spectrum_value(s.spectra[env["spectrum"]], s.knots, lam)
        b = host.spectrum_value(s.spectra[pt["spectrum"]], s.knots, lam)
        assert a == pytest.approx(b, rel=2e-6, abs=1e-7)
    v825, v830 = (host.spectrum_value(s.spectra[pt["spectrum"]], s.knots, l) for l in (825.0, 830.0))
    assert host.spectrum_value(s.spectra[env["spectrum"]], s.knots, 830.0) == pytest.approx(
        0.999 * v830 + 0.001 * v825, rel=1e-5)
    # a point light's power is spread over 4 pi steradians of a unit-luminance spectrum
    assert pt["scale"] == pytest.approx(2 / (4 * math.pi * spectrum_y(host, s, pt["spectrum"])), rel=3e-3)


def test_two_constant_environments_add(host, tmp_path):
    doc = {"lights": {"a": {"type": "environment", "color": [0.2, 0.4, 0.8], "intensity": 0.5},
                      "b": {"type": "environment", "intensity": 2.0}},
           "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
    f = tmp_path / "envs.prscene.json"
    f.write_text(json.dumps(doc))
    both = host.load_scene(f)
    env = both.lights[both.lights["kind"] == pa.LIGHT_ENV]
    assert len(env) == 1
    singles = []
    for k in ("a", "b"):
        d = dict(doc)
        d["lights"] = {k: doc["lights"][k]}
        g = tmp_path / f"env_{k}.prscene.json"
        g.write_text(json.dumps(d))
        singles.append(host.load_scene(g))
    for lam in (400.0, 475.0, 550.0, 700.0):
        total = sum(sc.lights[0]["scale"] * host.spectrum_value(sc.spectra[sc.lights[0]["spectrum"]], sc.knots, lam)
                    for sc in singles)
        got = env[0]["scale"] * host.spectrum_value(both.spectra[env[0]["spectrum"]], both.knots, lam)
        assert got == pytest.approx(total, rel=1e-5)


def test_spot_light_cones(host, studio):
    """cone_angle is the total width, falloff_start defaults to 0.8 of it and is capped at it, both in
    degrees, stored as FP32 cosines; the axis comes from direction or look_at (scene_loader.cpp:958-962,
    light.cpp:161-166)."""
    spot = studio.lights[studio.lights["kind"] == pa.LIGHT_SPOT][0]
    assert spot["cosFalloffEnd"] == np.float32(math.cos(math.radians(30)))
    assert spot["cosFalloffStart"] == pytest.approx(math.cos(math.radians(24)), rel=1e-6)
    assert np.allclose(spot["direction"], [0, -1, 0]) and np.allclose(spot["position"], [0, 3, 0])
    s = host.load_scene(os.path.join(HERE, "scenes", "spots", "scene.prscene.json"))
    spots = s.lights[s.lights["kind"] == pa.LIGHT_SPOT]
    hard = [l for l in spots if np.isclose(l["scale"], 35)][0]
    assert hard["cosFalloffStart"] == hard["cosFalloffEnd"] == np.float32(math.cos(math.radians(12)))
    d = np.array([-0.6, -1, -0.2])
    assert np.allclose(hard["direction"], d / np.linalg.norm(d), atol=1e-7)
    soft = [l for l in spots if np.isclose(l["scale"], 20)][0]
    t = np.array([0, 0, 0]) - np.array([-2, 3, 1])
    assert np.allclose(soft["direction"], t / np.linalg.norm(t), atol=1e-7)
    assert soft["cosFalloffStart"] == pytest.approx(math.cos(math.radians(30)), rel=1e-6)
    assert soft["cosFalloffEnd"] == pytest.approx(math.cos(math.radians(45)), rel=1e-6)


IES_SCENE = os.path.join(HERE, "scenes", "ies", "scene.prscene.json")


def ies_records(s):
    out = []
    for l in s.lights[s.lights["kind"] == pa.LIGHT_GONIO]:
        r = s.knots[l["profileOffset"]:l["profileOffset"] + l["profileSize"]]
        nV, nH = int(r[9]), int(r[10])
        assert l["profileSize"] == 12 + nV + nH + nV * nH
        out.append((l, r[:9].reshape(3, 3), nV, nH, r[11], r[12:12 + nV], r[12 + nV:12 + nV + nH],
                    r[12 + nV + nH:].reshape(nH, nV)))
    return out


def test_ies_records_frames_and_scales(host):
    """IES lights (scene_loader.cpp:940-955): the nadir follows direction / look_at (default down),
    'rotation' turns the frame about it, 'photometric' multiplies by the peak candela and 'power'
    divides by the profile's integral times the spectrum's luminance. Candela are stored relative to
    their peak after the multiplier and ballast factor; an INCLUDE tilt block is skipped."""
    s = host.load_scene(IES_SCENE)
    recs = {round(float(l["position"][0]), 3): (l, F, nV, nH, integ, V, H, C) for l, F, nV, nH, integ, V, H, C in ies_records(s)}
    assert len(recs) == 3 and not s.warnings
    for l, F, nV, nH, integ, V, H, C in recs.values():
        assert np.allclose(F @ F.T, np.eye(3), atol=1e-6) and np.linalg.det(F) > 0  # right-handed, orthonormal
        assert np.allclose(F[2], l["direction"]) and C.max() == 1 and C.min() >= 0
        assert np.all(np.diff(V) > 0) and np.all(np.diff(H) >= 0)
    down = recs[-1.5]
    assert np.allclose(down[1][2], [0, -1, 0]) and (down[2], down[3]) == (7, 3)
    assert down[0]["scale"] == pytest.approx(0.02)
    assert np.allclose(down[7][0], np.array([900, 880, 760, 520, 260, 80, 0]) / 900)
    aimed = recs[2.0]
    z = np.array([0, 0.5, 0]) - np.array([2.0, 2.5, 1.5])
    assert np.allclose(aimed[1][2], z / np.linalg.norm(z), atol=1e-6)
    assert (aimed[2], aimed[3]) == (10, 1)  # the INCLUDE tilt block was skipped
    assert aimed[0]["scale"] == pytest.approx(1.5e-5 * 1200 * 0.5 * 0.8, rel=1e-6)  # photometric: x peak cd
    # rotation: the frame's x axis is turned 30 degrees about z from the unrotated frame's x
    zz = aimed[1][2]
    sign = math.copysign(1, zz[2])
    a = -1 / (sign + zz[2])
    b = zz[0] * zz[1] * a
    fx = np.array([1 + sign * zz[0] ** 2 * a, sign * b, -sign * zz[0]])
    assert math.degrees(math.acos(np.clip(np.dot(aimed[1][0], fx), -1, 1))) == pytest.approx(30, abs=1e-3)
    washer = recs[0.5]
    y = spectrum_y(host, s, washer[0]["spectrum"])
    assert washer[0]["scale"] == pytest.approx(60 / (washer[4] * y), rel=3e-3)


def test_ies_integral_of_an_axial_profile(host):
    """For a rotationally symmetric profile I(theta) the integral over the sphere is
    2 pi int_0^pi I(theta) sin(theta) dtheta; the record's value (the reference's 128 x 64 x 4 cell
    estimate) must agree with fine quadrature of the piecewise-linear profile."""
    s = host.load_scene(IES_SCENE)
    l, F, nV, nH, integ, V, H, C = [r for r in ies_records(s) if r[3] == 1][0]
    th = np.linspace(0, 180, 200001)
    I = np.interp(th, V, C[0], left=0, right=0)
    exact = 2 * math.pi * np.trapezoid(I * np.sin(np.radians(th)), np.radians(th))
    assert integ == pytest.approx(exact, rel=2e-3)


def test_ies_parse_errors(host, tmp_path):
    def load(text):
        (tmp_path / "x.ies").write_text(text)
        doc = {"lights": {"g": {"type": "point", "position": [0, 2, 0], "ies": "x.ies"}},
               "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
        f = tmp_path / "g.prscene.json"
        f.write_text(json.dumps(doc))
        return host.load_scene(f)
    with pytest.raises(RuntimeError, match="missing TILT= line"):
        load("IESNA:LM-63-2002\n1 1 1 1 1 1 2 0 0 0\n")
    with pytest.raises(RuntimeError, match="only type C photometry"):
        load("TILT=NONE\n1 1 1 2 1 2 2 0 0 0\n1 1 1\n0 90\n0\n10 5\n")
    with pytest.raises(RuntimeError, match="truncated candela data"):
        load("TILT=NONE\n1 1 1 2 1 1 2 0 0 0\n1 1 1\n0 90\n0\n10\n")
    with pytest.raises(RuntimeError, match="all candela values are zero"):
        load("TILT=NONE\n1 1 1 2 1 1 2 0 0 0\n1 1 1\n0 90\n0\n0 0\n")
    ok = load("TILT=NONE\n1 1 2 2 1 1 2 0 0 0\n1 1 1\n0 90\n0\n10 5\n")
    rec = ies_records(ok)[0]
    assert np.allclose(rec[7][0], [1, 0.5])


def test_sun_light_and_its_spectrum_match_reference(host, tmp_path):
    """A sun light (scene_loader.cpp:977-997): direction from elevation / azimuth, disk radius
    0.2665 deg x sun_size, scale intensity x sun_scale, no light once the sun is 0.02 below the
    horizon; its radiance spectrum against the reference's MakeSunRadiance (sky.cpp:207-233: air
    mass, Rayleigh and aerosol depth, blackbody at 2e6 cd/m^2 scaled with the disk's size), golden
    vectors from the reference's own code. A sun-sky light adds its Preetham sky as an image
    environment of resolution x resolution / 2 (scene_loader.cpp:991-996)."""
    G = np.load(os.path.join(HERE, "golden", "ref_components.npz"))
    for (el, az, T, r), ref in zip(G["sun_cases"], G["sun_radiance"]):
        doc = {"lights": {"s": {"type": "sun", "elevation": float(el), "azimuth": float(az), "turbidity": float(T),
                                "sun_size": float(r) / 0.2665, "intensity": 0.5, "sun_scale": 2}},
               "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
        f = tmp_path / "sun.prscene.json"
        f.write_text(json.dumps(doc))
        s = host.load_scene(f)
        e, a = math.radians(el), math.radians(az)
        if math.sin(e) <= -0.02:
            assert len(s.lights) == 0 and not ref.any()
            continue
        sun = s.lights[0]
        assert sun["kind"] == pa.LIGHT_SUN and sun["scale"] == pytest.approx(1.0)
        assert np.allclose(sun["direction"], [math.cos(e) * math.sin(a), math.sin(e), math.cos(e) * math.cos(a)],
                           atol=1e-6)
        assert sun["cosFalloffEnd"] == pytest.approx(math.cos(math.radians(r)), abs=1e-7)
        got = np.array([host.spectrum_value(s.spectra[sun["spectrum"]], s.knots, float(l))
                        for l in np.arange(360.0, 831.0, 5.0)])
        np.testing.assert_allclose(got, ref, rtol=2e-4, atol=1e-30, err_msg=f"sun {el} {az} {T} {r}")
    doc["lights"]["s"].update(type="sun_sky", elevation=30.0, resolution=32)
    f.write_text(json.dumps(doc))
    s = host.load_scene(f)
    assert list(s.lights["kind"]) == [pa.LIGHT_SUN, pa.LIGHT_ENV]
    sky = s.lights[1]
    assert sky["profileSize"] > 0 and tuple(s.knots[sky["profileOffset"]:sky["profileOffset"] + 2]) == (32, 16)
    assert not any("Preetham sky" in w for w in s.warnings)


# ---- camera ----------------------------------------------------------------------------------------

def test_thin_lens_camera_from_f_stop_and_focus(studio, host, tmp_path):
    c = studio.camera
    assert c.tanHalfFov == pytest.approx(math.tan(math.radians(19)), rel=1e-6)
    focal_mm = 24 / (2 * math.tan(math.radians(19)))
    assert c.lensRadius == pytest.approx(focal_mm / 1000 / (2 * 4), rel=1e-5)
    assert c.focalDistance == pytest.approx(np.linalg.norm(np.array([0, 1.6, 4.2]) - [0, 0.35, 0.3]), rel=1e-6)
    m = np.array(c.cameraToWorld).reshape(3, 4)
    fwd = np.array([0, 0.35, 0]) - [0, 1.6, 4.2]
    assert np.allclose(m[:, 2], fwd / np.linalg.norm(fwd), atol=1e-6) and np.allclose(m[:, 3], [0, 1.6, 4.2])
    assert abs(np.dot(m[:, 0], [0, 1, 0])) < 1e-6  # right vector is horizontal
    # a pinhole with a focal length in millimetres
    doc = json.loads(re.sub(r"//[^\n]*|/\*.*?\*/", "", open(STUDIO).read(), flags=re.S))
    doc["render"]["camera"] = "alt"
    doc["include"] = [os.path.join(os.path.dirname(STUDIO), "materials.json")]
    doc["rigs"] = {}
    doc["active_rigs"] = []
    for g in doc["geometry"].values():
        if "file" in g:
            g["file"] = os.path.join(os.path.dirname(STUDIO), g["file"])
    f = tmp_path / "alt.prscene.json"
    f.write_text(json.dumps(doc))
    alt = host.load_scene(f)
    assert alt.camera.tanHalfFov == pytest.approx(24 / (2 * 35), rel=1e-5) and alt.camera.lensRadius == 0


# ---- meshes ----------------------------------------------------------------------------------------

def obj_reference(path):
    """The OBJ rules restated: 1-based or negative indices; one vertex per distinct (v, vt, vn) triple in
    order of first use; polygons fanned from their first vertex; uvs kept if any face names one (missing
    ones are 0); normals dropped altogether if any vertex lacks one."""
    P, T, N, verts, remap, tris = [], [], [], [], {}, []
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "v":
            P.append([float(x) for x in t[1:4]])
        elif t[0] == "vt":
            T.append([float(x) for x in t[1:3]])
        elif t[0] == "vn":
            N.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            face = []
            for w in t[1:]:
                parts = (w.split("/") + ["", ""])[:3]
                res = [(int(x) - 1 if int(x) > 0 else n + int(x)) if x else -1
                       for x, n in zip(parts, (len(P), len(T), len(N)))]
                key = tuple(res)
                if key not in remap:
                    remap[key] = len(verts)
                    verts.append(key)
                face.append(remap[key])
            tris += [(face[0], face[i], face[i + 1]) for i in range(1, len(face) - 1)]
    pos = np.array([P[v[0]] for v in verts])
    uv = np.array([T[v[1]] if v[1] >= 0 else [0, 0] for v in verts])
    has_n = all(v[2] >= 0 for v in verts)
    return pos, uv, has_n, tris


def single_object_scene(tmp_path, geom, transform=None):
    doc = {"geometry": {"g": geom}, "materials": {"m": {"type": "diffuse"}},
           "objects": [{"geometry": "g", "material": "m", **({"transform": transform} if transform else {})}],
           "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}
    f = tmp_path / "one.prscene.json"
    f.write_text(json.dumps(doc))
    return f


def test_obj_reader_matches_the_format_rules(host, tmp_path):
    obj = os.path.join(HERE, "scenes", "studio", "meshes", "gem.obj")
    s = host.load_scene(single_object_scene(tmp_path, {"type": "mesh", "file": obj}))
    pos, uv, has_n, tris = obj_reference(obj)
    assert len(s.prims) == len(tris) == 9 and not has_n
    for p, (a, b, c) in zip(s.prims, tris):
        assert np.array_equal(np.array([p["p0"], p["p1"], p["p2"]]), pos[[a, b, c]].astype(np.float32))
        assert np.array_equal(np.array([[p["uv"][0], p["uv"][1]], [p["uv"][2], p["uv"][3]], [p["uv"][4], p["uv"][5]]]),
                              uv[[a, b, c]].astype(np.float32))
        assert not np.any(p["n0"])  # a vertex without a normal drops shading normals for the mesh


def write_ply(path, verts, faces, binary, extra_element=False):
    head = ["ply", "format " + ("binary_little_endian" if binary else "ascii") + " 1.0",
            f"element vertex {len(verts)}", "property float x", "property float y", "property float z",
            "property float nx", "property float ny", "property float nz", "property float s", "property float t"]
    if extra_element:
        head += ["element junk 2", "property uchar a", "property list uchar short b"]
    head += [f"element face {len(faces)}", "property uchar flags", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if binary:
            for v in verts:
                f.write(struct.pack("<8f", *v))
            if extra_element:
                f.write(struct.pack("<BBhh", 7, 2, -3, 4) + struct.pack("<BB", 1, 0))
            for fc in faces:
                f.write(struct.pack("<BB", 0, len(fc)) + struct.pack(f"<{len(fc)}i", *fc))
        else:
            for v in verts:
                f.write((" ".join(repr(float(x)) for x in v) + "\n").encode())
            if extra_element:
                f.write(b"7 2 -3 4\n1 0\n")
            for fc in faces:
                f.write(("0 " + " ".join(str(x) for x in [len(fc)] + list(fc)) + "\n").encode())


def test_ply_ascii_and_binary_agree_and_normalise_normals(host, tmp_path):
    verts = [(0, 0, 0, 0, 0, 2, 0, 0), (1, 0, 0, 0, 0, 3, 1, 0), (1, 1, 0, 0, 0, 1, 1, 1), (0, 1, 0, 0, 0, 5, 0, 1),
             (0.5, 0.5, 1, 0, 0, 1, 0.5, 0.5)]
    faces = [(0, 1, 2, 3), (0, 1, 4)]
    loaded = []
    for binary in (False, True):
        p = tmp_path / f"m{int(binary)}.ply"
        write_ply(p, verts, faces, binary, extra_element=True)
        loaded.append(host.load_scene(single_object_scene(tmp_path, {"type": "mesh", "file": str(p)})))
    a, b = loaded
    assert len(a.prims) == 3 and np.array_equal(a.prims.view(np.uint8), b.prims.view(np.uint8))
    assert np.allclose(a.prims["n0"], [0, 0, 1]) and np.allclose(a.prims["n2"], [0, 0, 1])
    assert np.allclose(a.prims[1]["p2"], [0, 1, 0]) and np.allclose(a.prims[1]["uv"][4:6], [0, 1])


def test_mirroring_transform_keeps_outward_winding(host, tmp_path):
    wedge = os.path.join(HERE, "scenes", "studio", "meshes", "wedge.ply")
    plain = host.load_scene(single_object_scene(tmp_path, {"type": "mesh", "file": wedge}))
    mirror = host.load_scene(single_object_scene(tmp_path, {"type": "mesh", "file": wedge},
                                                 {"matrix": [-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]}))
    for p, q in zip(plain.prims, mirror.prims):
        n, _ = tri_normal(p)
        m, _ = tri_normal(q)
        assert np.allclose(m, n * [-1, 1, 1], atol=1e-6)  # the mirror image of the outward normal


# ---- errors ----------------------------------------------------------------------------------------

BASE = {"materials": {"m": {"type": "diffuse"}}, "geometry": {"g": {"type": "sphere"}},
        "objects": [{"geometry": "g", "material": "m"}], "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}


@pytest.mark.parametrize("patch,message", [
    ({"objects": [{"geometry": "g", "material": "nope"}]}, "unknown material 'nope'"),
    ({"objects": [{"geometry": "nope", "material": "m"}]}, "unknown geometry 'nope'"),
    ({"objects": [{"geometry": "g"}]}, "missing 'material'"),
    ({"active_rigs": ["nope"]}, "unknown light rig 'nope'"),
    ({"materials": {"m": {"type": "diffuse", "reflectance": {"texture": "t"}}},
      "textures": {"t": {"type": "image", "file": "x.png"}}}, "texture 't': cannot open"),
    ({"materials": {"m": {"type": "diffuse", "reflectance": {"texture": "t"}}},
      "textures": {"t": {"type": "noise"}}}, "noise colour textures are not supported"),
    ({"materials": {"m": {"type": "plastic"}}}, "unknown type 'plastic'"),
    ({"materials": {"m": {"type": "conductor", "metal": "Pt"}}}, "unknown metal 'Pt'"),
    ({"materials": {"m": {"type": "dielectric", "ior": "unobtainium"}}}, "unknown IOR 'unobtainium'"),
    ({"cameras": {"c": {"type": "orthographic", "position": [0, 1, 5], "look_at": [0, 0, 0]}}}, "perspective/thin-lens"),
    ({"render": {"camera": "other"}}, "camera 'other' not found"),
    ({"geometry": {"g": {"type": "mesh", "file": "missing.obj"}}}, "cannot open"),
    ({"geometry": {"g": {"type": "gltf", "file": "x.gltf"}}}, "glTF meshes are not supported"),
    ({"lights": {"l": {"type": "laser"}}}, "light 'l': unknown type 'laser'"),
    ({"media": {"h": {"type": "noise"}}, "objects": [{"geometry": "g", "material": "m", "interior": "h"}]},
     "noise media are not supported"),
    ({"render": {"film": {"filter": "lanczos"}}}, "unknown filter 'lanczos'"),
    ({"materials": {"m": {"type": "diffuse", "reflectance": [1, 2]}}}, "expected [x, y, z]"),
])
def test_load_errors(host, tmp_path, patch, message):
    doc = json.loads(json.dumps(BASE))
    doc.update(patch)
    f = tmp_path / "bad.prscene.json"
    f.write_text(json.dumps(doc))
    with pytest.raises(RuntimeError) as e:
        host.load_scene(f)
    assert message in str(e.value)


def test_syntax_errors_and_include_cycles(host, tmp_path):
    f = tmp_path / "syntax.prscene.json"
    f.write_text('{\n  // a comment\n  "objects": [\n    {"geometry": "g",, }\n  ]\n}\n')
    with pytest.raises(RuntimeError, match=r"syntax\.prscene\.json:4: "):
        host.load_scene(f)
    g = tmp_path / "loop.prscene.json"
    g.write_text('{"include": ["loop.prscene.json"]}')
    with pytest.raises(RuntimeError, match="include depth exceeded"):
        host.load_scene(g)
    with pytest.raises(RuntimeError, match="cannot open"):
        host.load_scene(tmp_path / "absent.prscene.json")


def test_comments_escapes_and_unicode_names(host, tmp_path):
    f = tmp_path / "c.prscene.json"
    f.write_text('/* block\n comment */ {"materials": {"m\\u00e9": {"type": "diffuse"}}, // line\n'
                 ' "geometry": {"g": {"type": "sphere", "radius": 2e-1}},\n'
                 ' "objects": [{"geometry": "g", "material": "m\\u00e9", "name": "tab\\tname"}],\n'
                 ' "cameras": {"c": {"position": [0, 1, 5], "look_at": [0, 0, 0]}}}')
    s = host.load_scene(f)
    assert len(s.prims) == 1 and s.prims[0]["r"] == pytest.approx(0.2)


# ---- the reference's own scene files (read where the reference tree is present) --------------------

def strip_comments(text):
    return re.sub(r"//[^\n]*|/\*.*?\*/", "", text, flags=re.S)


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference scene files not present")
def test_reference_sphere_pyramid(host):
    path = os.path.join(REF_SCENES, "sphere-pyramid", "scene.prscene.json")
    s = host.load_scene(path)
    doc = json.loads(strip_comments(open(path).read()))
    assert len(sphere_prims(s)) == sum(1 for o in doc["objects"] if o["geometry"] == "ball") == 14
    assert len(s.prims) == 2 + 14 + 2 + 2  # floor, balls, key and fill rects
    assert sorted(s.lights["kind"].tolist()) == sorted([pa.LIGHT_AREA] * 5 + [pa.LIGHT_ENV])
    assert (s.width, s.height) == (960, 540) and s.defaults["maxDepth"] == 64 and s.defaults["seed"] == 1
    fog = host.load_scene(path, rigs=["studio", "fog"])
    assert fog.desc.cameraMedium >= 0 and len(fog.prims) == len(s.prims) + 12
    day = host.load_scene(path, rigs=["daylight"])
    # the sun_sky rig: its sun and its Preetham sky map beside the glowing ball
    assert sorted(day.lights["kind"].tolist()) == sorted([pa.LIGHT_AREA, pa.LIGHT_SUN, pa.LIGHT_ENV])
    assert not any("Preetham sky" in w for w in day.warnings)


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason="reference scene files not present")
def test_reference_diamonds(host):
    root = os.path.join(REF_SCENES, "diamonds")
    s = host.load_scene(os.path.join(root, "scene.prscene.json"))
    doc = json.loads(strip_comments(open(os.path.join(root, "scene.prscene.json")).read()))

    def tri_count(fname):
        p = os.path.join(root, fname)
        if p.endswith(".obj"):
            return sum(len(l.split()) - 3 for l in open(p) if l.startswith("f "))
        head = open(p, "rb").read(2048).split(b"end_header")[0].decode()
        return int(re.search(r"element face (\d+)", head).group(1))  # triangles only in this file

    expected = sum(tri_count(doc["geometry"][o["geometry"]]["file"]) for o in doc["objects"])
    rig = json.loads(strip_comments(open(os.path.join(root, "rigs", "jeweller.json")).read()))
    rects = sum(1 for l in rig["lights"].values() if l["type"] == "rect")
    assert len(s.prims) == expected + 2 * rects
    # the silk's bump map is the scene's 512x512 8-bit grey PNG, read as linear data, tiled 35 times
    shaded = [int(m["shading"]) for m in s.materials if m["shading"] > 0]
    assert len(shaded) == 1
    rec = s.knots[shaded[0] - 1:shaded[0] + 7]
    assert (rec[0], rec[4], rec[6]) == (0, 1, -1) and rec[7] == pytest.approx(0.01)
    bump = s.knots[int(rec[5]):]
    np.testing.assert_array_equal(bump[:8], [512, 512, 0, 1, 35, 35, 0, 0])
    texels = bump[12:12 + 512 * 512 * 3].reshape(512, 512, 3)
    np.testing.assert_array_equal(texels[..., 0], texels[..., 2])
    np.testing.assert_array_equal(np.round(texels * 255) / np.float32(255), texels)
    assert 0.1 < texels.mean() < 0.9 and texels.min() == 0 and texels.max() == 1
    assert not any("bump" in w for w in s.warnings)
    both = host.load_scene(os.path.join(root, "scene.prscene.json"), rigs=["sparkle", "window"])
    assert (both.lights["kind"] == pa.LIGHT_ENV).sum() == 1
