"""Image textures, normal maps and bump maps on the CPU: the loader's records, and what the
mathematics fixes for the oracle's path loop with them.

The texture lookup and the shading perturbation themselves are pinned bit for bit against the
reference's compiled ImageData::Lookup and Material::PerturbShading (test_oracle_golden.py). Here:
- the loader writes an image texture (`"type": "image"`: PFM / PPM, colorspace, wrap, flip_v,
  uv_scale, uv_offset, multiplier) as a SpecImage reflectance with its record, and a material's
  normal_map / bump_map as its shading record (scene_loader.cpp:438-553, 732-752);
- under an overhead distant light and direct lighting only, a Lambertian plane's radiance scales
  with the cosine between the light and the shading normal. A constant normal map t = 2c - 1 gives
  the shading normal's height t_z / |t| (independent of the tangent frame), and a height that rises
  linearly in u with slope s (in height per unit u) on a plane whose dpdu has length L_u tilts it
  to L_u / sqrt(s^2 + L_u^2): each pixel of the perturbed render equals the plain render times that
  factor, sample for sample;
- a reflectance-1 sphere with a white image texture under a white environment is a white furnace;
  a constant-colour texture seen under it has the sRGB -> XYZ of its colour within 2%.
"""
import os

import numpy as np
import pytest

import envscenes as es
import oracle as orc
import pathscenes as ps
import prender_amd as pa
import texscenes as ts

REF_SILK = "/root/reference/scenes/diamonds/textures/silk_weave.png"


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("textures"))


def _render(scene, spp=4, depth=1, seed=3):
    s, _, _ = orc.OracleScene(scene).render(pa.render_params(spp=spp, first_sample=0, seed=seed, max_depth=depth),
                                            variance=False)
    return s


def test_image_texture_records(host, tmpdir):
    img = np.random.default_rng(2).random((3, 5, 3)).astype(np.float32)
    es.write_pfm(f"{tmpdir}/a.pfm", img)
    rgb8 = (np.random.default_rng(3).random((2, 4, 3)) * 255).astype(np.uint8)
    es.write_ppm(f"{tmpdir}/b.ppm", rgb8)
    doc = {"textures": {"a": {"type": "image", "file": f"{tmpdir}/a.pfm", "wrap": "mirror", "flip_v": False,
                              "uv_scale": [2, 3], "uv_offset": [0.25, -0.5], "multiplier": [0.5, 1, 2]},
                        "a2": {"type": "image", "file": f"{tmpdir}/a.pfm", "wrap": "mirror", "flip_v": False,
                               "uv_scale": [2, 3], "uv_offset": [0.25, -0.5], "multiplier": [0.5, 1, 2]},
                        "b": {"type": "image", "file": f"{tmpdir}/b.ppm"},
                        "bl": {"type": "image", "file": f"{tmpdir}/b.ppm", "colorspace": "linear"}},
           "materials": {"m1": {"type": "diffuse", "reflectance": {"texture": "a"}},
                         "m2": {"type": "conductor", "reflectance": {"texture": "a2"}, "roughness": 0.2},
                         "m3": {"type": "diffuse", "reflectance": {"texture": "b"},
                                "normal_map": {"texture": "bl", "strength": 2, "flip_green": True},
                                "bump_map": {"texture": "b", "channel": "g", "scale": 0.05}}},
           "geometry": {"r": {"type": "rect"}},
           "objects": [{"geometry": "r", "material": m} for m in ("m1", "m2", "m3")],
           "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    s = es.load(host, tmpdir, "records", doc)
    mats = s.materials
    a = s.spectra[mats[0]["albedo"]]
    assert a["kind"] == pa.SPEC_IMAGE
    assert mats[1]["albedo"] == mats[0]["albedo"] or s.spectra[mats[1]["albedo"]]["offset"] == a["offset"]
    rec = s.knots[a["offset"]:a["offset"] + 12 + 45]
    np.testing.assert_array_equal(rec[:11], [5, 3, 2, 0, 2, 3, 0.25, -0.5, 0.5, 1, 2])
    np.testing.assert_array_equal(rec[12:], img.reshape(-1))
    table = int(rec[11])
    assert s.knots[table] == 64
    b = s.knots[s.spectra[mats[2]["albedo"]]["offset"]:][:12 + 24]
    v = rgb8.reshape(-1) / 255.0
    np.testing.assert_allclose(b[12:], np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4), rtol=1e-5,
                               atol=1e-7)
    assert b[3] == 1  # flip_v defaults to true
    sh = s.knots[mats[2]["shading"] - 1:][:8]
    assert (sh[0], sh[2], sh[3], sh[4], sh[6]) == (1, 2, 1, 1, 1) and sh[7] == pytest.approx(0.05)
    nrec = s.knots[int(sh[1]):][:12 + 24]
    np.testing.assert_allclose(nrec[12:], v, rtol=1e-6)  # normal maps are read as data
    assert mats[0]["shading"] == 0 and mats[1]["shading"] == 0


def test_textures_outside_reflectance_are_named_errors(host, tmpdir):
    es.write_pfm(f"{tmpdir}/c.pfm", es.constant_map((0.5, 0.5, 0.5), 2, 2))
    base = {"textures": {"c": {"type": "image", "file": f"{tmpdir}/c.pfm"},
                         "chk": {"type": "checker", "a": {"texture": "c"}, "b": 0.2}},
            "geometry": {"r": {"type": "rect"}}, "objects": [{"geometry": "r", "material": "m"}],
            "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    for m, msg in (({"type": "diffuse", "sheen": {"color": {"texture": "c"}}}, "reflectance only"),
                   ({"type": "coated_diffuse", "albedo": {"texture": "c"}}, "reflectance only"),
                   ({"type": "diffuse", "reflectance": {"texture": "chk"}}, "reflectance only"),
                   ({"type": "diffuse", "normal_map": {"texture": "chk"}}, "normal map is an image")):
        doc = dict(base, materials={"m": m})
        with pytest.raises(RuntimeError, match=msg):
            es.load(host, tmpdir, "bad", doc)


def test_normal_map_scales_direct_light_by_the_shading_cosine(host, tmpdir):
    c = np.array([0.8, 0.5, 0.7], np.float32)
    t = 2 * c - 1
    want = t[2] / np.linalg.norm(t)
    plain = _render(es.load(host, tmpdir, "np0", ts.plane_doc(tmpdir)))
    mapped = _render(es.load(host, tmpdir, "np1", ts.plane_doc(tmpdir, normal=es.constant_map(c, 3, 3))))
    assert plain.min() > 0
    ratio = mapped / plain
    print(f"normal map: pixel ratio {ratio.min():.6f}..{ratio.max():.6f}, closed form {want:.6f}", flush=True)
    np.testing.assert_allclose(ratio, want, rtol=2e-5)


def test_bump_ramp_tilts_the_shading_normal(host, tmpdir):
    """A 2-texel ramp (0 -> 1) with clamp wrap is linear for u between the texel centres 0.25 and
    0.75 with slope 2; the bump scale 4 makes s = 8 against |dpdu| = 16 (the rect's width)."""
    size, scale = 16.0, 4.0
    want = size / np.hypot(2 * scale, size)
    plain = _render(es.load(host, tmpdir, "bp0", ts.plane_doc(tmpdir, size=size)))
    bumped = _render(es.load(host, tmpdir, "bp1", ts.plane_doc(tmpdir, size=size, bump=(ts.ramp(2), scale))))
    ratio = bumped / plain
    print(f"bump ramp: pixel ratio {ratio.min():.6f}..{ratio.max():.6f}, closed form {want:.6f}", flush=True)
    np.testing.assert_allclose(ratio, want, rtol=2e-3)


def test_white_texture_furnace_and_coloured_texture(host, tmpdir):
    es.write_pfm(f"{tmpdir}/white_env.pfm", es.constant_map((1, 1, 1)))
    doc = es.furnace_env_doc(f"{tmpdir}/white_env.pfm")
    es.write_pfm(f"{tmpdir}/tex_white.pfm", es.constant_map((1, 1, 1), 4, 4))
    doc["textures"] = {"w": {"type": "image", "file": f"{tmpdir}/tex_white.pfm"}}
    doc["materials"]["m"]["reflectance"] = {"texture": "w"}
    s = es.load(host, tmpdir, "tex_furnace", doc)
    spp = 128
    su, sq, _ = orc.OracleScene(s).render(pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=-1))
    mean, se = ps.mean_and_se(su, sq, spp)
    print(f"white-texture furnace: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1) < 0.01
    for k in (0, 2):
        assert abs(mean[k] / mean[1] / ps.D65_XYZ[k] - 1) < 0.01
    # a constant-colour texture on a sphere seen under the white map, one bounce (a convex sphere
    # does not see itself): radiance = albedo x D65
    rgb = (0.7, 0.35, 0.15)
    es.write_pfm(f"{tmpdir}/tex_col.pfm", es.constant_map(rgb, 4, 4))
    doc["textures"] = {"w": {"type": "image", "file": f"{tmpdir}/tex_col.pfm"}}
    doc["cameras"]["cam"]["fov_y"] = 12
    s = es.load(host, tmpdir, "tex_colour", doc)
    su, sq, _ = orc.OracleScene(s).render(pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=-1))
    mean, se = ps.mean_and_se(su, sq, spp)
    srgb_to_xyz = np.array([[0.4124564, 0.3575761, 0.1804375], [0.2126729, 0.7151522, 0.0721750],
                            [0.0193339, 0.1191920, 0.9503041]])
    want = srgb_to_xyz @ np.array(rgb)
    print(f"colour texture: XYZ = {mean.round(5)} (sRGB->XYZ {want.round(5)})", flush=True)
    np.testing.assert_allclose(mean, want, rtol=0.02)


PNG_KINDS = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8),
             (4, 16), (6, 8), (6, 16)]
PNG_FILTERS = [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4, 4, 3)]


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("colour_type,depth", PNG_KINDS)
def test_png_textures_decode_to_their_samples(host, tmpdir, colour_type, depth, interlace):
    """PNG images written here by the specification's definitions (every colour type and bit depth,
    each row filter alone and mixed, plain and Adam7, odd sizes so that sub-byte rows are padded and
    interlace passes are partial or empty) load as linear texels equal to their samples reduced to
    8 bits over 255. Filter 0 fixes the layout independently of any predictor; a decoder predictor
    that differs from the specification's fails the other filters."""
    rng = np.random.default_rng(colour_type * 100 + depth + 7 * interlace)
    nc = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[colour_type]
    palette = None
    textures, mats, objects, want = {}, {}, [], {}
    for k, filters in enumerate(PNG_FILTERS):
        h, w = [(11, 13), (1, 1), (3, 2), (9, 17), (5, 3), (16, 7)][k]
        hi = 2 ** depth
        if colour_type == 3:
            n_pal = min(hi, 11)
            palette = rng.integers(0, 256, (n_pal, 3))
            samples = rng.integers(0, n_pal, (h, w, 1))
        else:
            samples = rng.integers(0, hi, (h, w, nc))
        name = f"p{colour_type}_{depth}_{int(interlace)}_{k}"
        ts.write_png(f"{tmpdir}/{name}.png", samples, colour_type, depth, filters, interlace, palette)
        textures[name] = {"type": "image", "file": f"{tmpdir}/{name}.png", "colorspace": "linear"}
        mats[name] = {"type": "diffuse", "reflectance": {"texture": name}}
        objects.append({"geometry": "r", "material": name})
        want[name] = (ts.png_expected_rgb8(samples, colour_type, depth, palette), w, h)
    doc = {"textures": textures, "materials": mats, "geometry": {"r": {"type": "rect"}}, "objects": objects,
           "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    s = es.load(host, tmpdir, f"png{colour_type}_{depth}_{int(interlace)}", doc)
    for i, name in enumerate(mats):
        rgb8, w, h = want[name]
        off = s.spectra[s.materials[i]["albedo"]]["offset"]
        assert s.spectra[s.materials[i]["albedo"]]["kind"] == pa.SPEC_IMAGE
        assert (s.knots[off], s.knots[off + 1]) == (w, h)
        texels = s.knots[off + 12:off + 12 + w * h * 3]
        np.testing.assert_array_equal(texels, rgb8.reshape(-1).astype(np.float32) / np.float32(255), err_msg=name)


def test_png_srgb_decode_and_errors(host, tmpdir):
    rgb8 = np.random.default_rng(9).integers(0, 256, (4, 6, 3))
    ts.write_png(f"{tmpdir}/s.png", rgb8, 2, 8, (4,))
    doc = {"textures": {"t": {"type": "image", "file": f"{tmpdir}/s.png"}},
           "materials": {"m": {"type": "diffuse", "reflectance": {"texture": "t"}}},
           "geometry": {"r": {"type": "rect"}}, "objects": [{"geometry": "r", "material": "m"}],
           "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    s = es.load(host, tmpdir, "png_srgb", doc)
    off = s.spectra[s.materials[0]["albedo"]]["offset"]
    v = rgb8.reshape(-1) / 255.0
    np.testing.assert_allclose(s.knots[off + 12:off + 12 + 72],
                               np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4), rtol=1e-5, atol=1e-7)
    good = open(f"{tmpdir}/s.png", "rb").read()
    for name, data, msg in (("trunc", good[:60], "truncated PNG"),
                            ("nosig", b"PNG" + good[3:], "not a PNG file"),
                            ("idat", good[:33] + good[33:].replace(b"IDAT", b"IDAx", 1), "unsupported PNG chunk")):
        with open(f"{tmpdir}/{name}.png", "wb") as f:
            f.write(data)
        d = dict(doc, textures={"t": {"type": "image", "file": f"{tmpdir}/{name}.png"}})
        with pytest.raises(RuntimeError, match=msg):
            es.load(host, tmpdir, "png_" + name, d)


def test_png_decoding_equals_the_reference_decoder(host, tmpdir):
    """The reference reads PNG through its vendored stb_image asking for 8-bit RGB
    (src/core/image.cpp:90): the golden vectors hold that decoder's output for PNG files of every
    colour type and bit depth, plain and interlaced, with every row filter. The loader's linear
    texels times 255 are those bytes exactly."""
    G = np.load(f"{os.path.dirname(os.path.abspath(__file__))}/golden/ref_components.npz")
    ends, shapes = G["png_file_ends"], G["png_shapes"]
    starts = np.concatenate([[0], ends[:-1]])
    tstart = 0
    textures, mats, objects, want = {}, {}, [], []
    for k, (a, b) in enumerate(zip(starts, ends)):
        path = f"{tmpdir}/golden{k}.png"
        with open(path, "wb") as f:
            f.write(G["png_file_bytes"][a:b].tobytes())
        h, w = shapes[k]
        want.append(G["png_texels"][tstart:tstart + h * w * 3])
        tstart += h * w * 3
        textures[f"t{k}"] = {"type": "image", "file": path, "colorspace": "linear"}
        mats[f"m{k}"] = {"type": "diffuse", "reflectance": {"texture": f"t{k}"}}
        objects.append({"geometry": "r", "material": f"m{k}"})
    doc = {"textures": textures, "materials": mats, "geometry": {"r": {"type": "rect"}}, "objects": objects,
           "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    s = es.load(host, tmpdir, "png_golden", doc)
    for k, wk in enumerate(want):
        off = s.spectra[s.materials[k]["albedo"]]["offset"]
        texels = s.knots[off + 12:off + 12 + wk.size]
        np.testing.assert_array_equal(np.round(texels * 255).astype(np.uint8), wk, err_msg=f"file {k}")
        np.testing.assert_array_equal(texels, wk.astype(np.float32) / np.float32(255))


@pytest.mark.skipif(not os.path.exists(REF_SILK), reason="the reference tree is not present")
def test_reference_silk_texture_decodes_as_the_reference_decoder(host, tmpdir):
    import hashlib
    G = np.load(f"{os.path.dirname(os.path.abspath(__file__))}/golden/ref_components.npz")
    doc = {"textures": {"t": {"type": "image", "file": REF_SILK, "colorspace": "linear"}},
           "materials": {"m": {"type": "diffuse", "reflectance": {"texture": "t"}}},
           "geometry": {"r": {"type": "rect"}}, "objects": [{"geometry": "r", "material": "m"}],
           "cameras": {"cam": {"position": [0, 2, 3], "look_at": [0, 0, 0]}}}
    s = es.load(host, tmpdir, "silk", doc)
    h, w = G["silk_shape"]
    off = s.spectra[s.materials[0]["albedo"]]["offset"]
    texels = np.round(s.knots[off + 12:off + 12 + h * w * 3] * 255).astype(np.uint8)
    assert hashlib.sha256(texels.tobytes()).digest() == G["silk_sha256"].tobytes()
