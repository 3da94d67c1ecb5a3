"""Image environment lights (src/lights/light.cpp:215-297) and the sun-sky's Preetham map, on the
CPU: the loader's PFM / PPM / PNG / Radiance HDR / OpenEXR reading and its record in the scene's knots; the oracle's sampling
against the host's Distribution2D tables (themselves pinned bit for bit against the reference's);
and what the mathematics fixes for the path loop under an image environment:
- a pixel that sees a constant map directly gets the map's colour as an RGB illuminant: for white,
  exactly the D65 white at Y = 1 (the uplift's grey closed form); for a colour, the sRGB-to-XYZ of
  that colour within 1% (the sigmoid spectrum fitted to that sRGB, through the trilinear table);
- a reflectance-1 diffuse sphere under a constant white map is a white furnace: mean Y = 1 within
  1%, which pins the map's importance sampling and its pdf in the MIS weights.
"""
import os

import numpy as np
import pytest

import envscenes as es
import oracle as orc
import pathscenes as ps
import prender_amd as pa
import texscenes as ts

# linear sRGB -> XYZ (IEC 61966-2-1, D65)
SRGB_TO_XYZ = np.array([[0.4124564, 0.3575761, 0.1804375], [0.2126729, 0.7151522, 0.0721750],
                        [0.0193339, 0.1191920, 0.9503041]])


@pytest.fixture(scope="module")
def tmpdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("envmap"))


def _env_light(scene):
    env = np.flatnonzero((scene.lights["kind"] == pa.LIGHT_ENV) & (scene.lights["profileSize"] > 0))
    assert len(env) == 1
    return int(env[0])


def test_pfm_and_ppm_maps_load_into_the_record(host, tmpdir):
    img = es.studio_map(32, 16)
    es.write_pfm(f"{tmpdir}/m.pfm", img)
    s = es.load(host, tmpdir, "pfm", es.env_doc(f"{tmpdir}/m.pfm", objects=False, rotation=30, intensity=2.0))
    li = _env_light(s)
    L = s.lights[li]
    rec = s.knots[L["profileOffset"]:L["profileOffset"] + L["profileSize"]]
    assert (rec[0], rec[1]) == (32, 16)
    np.testing.assert_allclose([rec[2], rec[3]], [np.cos(np.radians(30)), np.sin(np.radians(30))], rtol=1e-6)
    assert L["scale"] == pytest.approx(2.0)
    np.testing.assert_array_equal(rec[8:8 + 32 * 16 * 3], img.reshape(-1))  # top row first
    assert L["profileSize"] == 8 + 32 * 16 * 3 + 32 * 16 + 16 * 33 + 16 + 17
    table = s.knots[int(rec[7]):]
    assert table[0] == 64
    # the record's sampling tables are the host's Distribution2D of the map
    u = np.random.default_rng(1).random((64, 2)).astype(np.float32)
    out, avg = host.envmap_sample(img, u)
    assert rec[5] == pytest.approx(avg, rel=1e-6)
    # PPM: 8-bit sRGB, decoded
    rgb8 = (np.random.default_rng(2).random((4, 5, 3)) * 255).astype(np.uint8)
    es.write_ppm(f"{tmpdir}/m.ppm", rgb8)
    s2 = es.load(host, tmpdir, "ppm", es.env_doc(f"{tmpdir}/m.ppm", objects=False))
    L2 = s2.lights[_env_light(s2)]
    r2 = s2.knots[L2["profileOffset"] + 8:L2["profileOffset"] + 8 + 60]
    v = rgb8.reshape(-1) / 255.0
    lin = np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(r2, lin, rtol=1e-5, atol=1e-7)


def test_unreadable_maps_are_named_errors(host, tmpdir):
    with open(f"{tmpdir}/bad.jpg", "wb") as f:
        f.write(b"\xff\xd8\xff....")
    with pytest.raises(RuntimeError, match="not read by this loader"):
        es.load(host, tmpdir, "jpg", es.env_doc(f"{tmpdir}/bad.jpg", objects=False))
    with open(f"{tmpdir}/bad.png", "wb") as f:
        f.write(b"\x89PNG....")
    with pytest.raises(RuntimeError, match="not a PNG file"):
        es.load(host, tmpdir, "png", es.env_doc(f"{tmpdir}/bad.png", objects=False))
    with open(f"{tmpdir}/short.pfm", "wb") as f:
        f.write(b"PF\n4 4\n-1.0\n" + b"\0" * 20)
    with pytest.raises(RuntimeError, match="truncated PFM"):
        es.load(host, tmpdir, "short", es.env_doc(f"{tmpdir}/short.pfm", objects=False))


def test_oracle_env_sampling_matches_the_host_tables(host, tmpdir):
    """The oracle rebuilds the map's Distribution2D from the texels; its SampleLi direction and pdf
    must be the host tables' sample (pinned against the reference) mapped through the rotation."""
    img = es.studio_map(48, 24)
    es.write_pfm(f"{tmpdir}/s.pfm", img)
    rot = 25.0
    s = es.load(host, tmpdir, "samp", es.env_doc(f"{tmpdir}/s.pfm", objects=True, rotation=rot))
    li = _env_light(s)
    o = orc.OracleScene(s)
    rng = np.random.default_rng(4)
    U = rng.random((400, 2)).astype(np.float32)
    ref, _ = host.envmap_sample(img, U)
    L = orc.lib()
    c, sn = np.cos(np.radians(rot)), np.sin(np.radians(rot))
    p = orc.f32(0.1, 0.2, 0.3)
    n = orc.f32(0, 1, 0)
    lam, lpdf = orc.f32(450, 520, 610, 700), orc.f32(1, 1, 1, 1)
    out = np.zeros(15, np.float32)
    for i in range(len(U)):
        L.or_light_sample_li(o.h, li, p.ctypes.data, n.ctypes.data, orc.f32(*U[i]).ctypes.data, lam.ctypes.data,
                             lpdf.ctypes.data, out.ctypes.data)
        uu, vv, pdf, _ = ref[i]
        st = np.sin(vv * np.pi)
        if out[14] == 0:
            assert pdf == 0 or st == 0
            continue
        theta, phi = vv * np.pi, uu * 2 * np.pi
        dl = np.array([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)])
        dw = np.array([c * dl[0] + sn * dl[2], dl[1], -sn * dl[0] + c * dl[2]])  # Rotate(rot, +y)
        np.testing.assert_allclose(out[4:7], dw, atol=2e-5)
        np.testing.assert_allclose(out[7], pdf / (2 * np.pi * np.pi * st), rtol=2e-5)
        # PDF_Li of the sampled direction is the sampling pdf (same cell)
        q = L.or_light_pdf_li(o.h, li, p.ctypes.data, orc.f32(*out[4:7]).ctypes.data)
        assert q == pytest.approx(out[7], rel=1e-3)


def _direct_view_xyz(host, tmpdir, name, rgb):
    es.write_pfm(f"{tmpdir}/{name}.pfm", es.constant_map(rgb))
    s = es.load(host, tmpdir, name, es.env_doc(f"{tmpdir}/{name}.pfm", objects=False, film=(8, 8)))
    spp = 64
    su, sq, _ = orc.OracleScene(s).render(pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=-1))
    return ps.mean_and_se(su, sq, spp)


def test_white_map_seen_directly_is_d65_at_unit_luminance(host, tmpdir):
    mean, se = _direct_view_xyz(host, tmpdir, "white", (1, 1, 1))
    print(f"white map: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1) < max(0.01, 5 * se[1])
    for c in (0, 2):
        assert abs(mean[c] / mean[1] / ps.D65_XYZ[c] - 1) < 0.01


@pytest.mark.parametrize("rgb", [(0.8, 0.3, 0.1), (0.1, 0.5, 0.9), (2.0, 1.5, 0.25)])
def test_coloured_map_seen_directly_has_its_colour(host, tmpdir, rgb):
    mean, se = _direct_view_xyz(host, tmpdir, "c%d" % int(1000 * sum(rgb)), rgb)
    want = SRGB_TO_XYZ @ np.array(rgb)
    print(f"map {rgb}: XYZ = {mean.round(5)} (sRGB->XYZ {want.round(5)})", flush=True)
    np.testing.assert_allclose(mean, want, rtol=0.02)


def test_white_map_furnace(host, tmpdir):
    es.write_pfm(f"{tmpdir}/wf.pfm", es.constant_map((1, 1, 1)))
    s = es.load(host, tmpdir, "wfurnace", es.furnace_env_doc(f"{tmpdir}/wf.pfm"))
    spp = 128
    su, sq, _ = orc.OracleScene(s).render(pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=-1))
    mean, se = ps.mean_and_se(su, sq, spp)
    print(f"image-map furnace: XYZ = {mean.round(5)} +- {se.round(5)}", flush=True)
    assert abs(mean[1] - 1) < 0.01
    for c in (0, 2):
        assert abs(mean[c] / mean[1] / ps.D65_XYZ[c] - 1) < 0.01


def test_sun_sky_adds_the_preetham_map(host, tmpdir):
    s = es.load(host, tmpdir, "sunsky", es.sun_sky_doc(resolution=64))
    kinds = list(s.lights["kind"])
    assert kinds.count(pa.LIGHT_SUN) == 1
    li = _env_light(s)
    rec = s.knots[s.lights[li]["profileOffset"]:]
    assert (rec[0], rec[1]) == (64, 32)
    sun = s.lights[kinds.index(pa.LIGHT_SUN)]["direction"]
    sky = host.preetham_sky(sun, 3.0, 0.3, 64, 32)
    np.testing.assert_array_equal(rec[8:8 + 64 * 32 * 3], sky.reshape(-1))
    assert not any("sky" in w and "skipped" in w for w in s.warnings)


def _map_texels(host, tmpdir, name, path):
    s = es.load(host, tmpdir, name, es.env_doc(path, objects=False))
    L = s.lights[_env_light(s)]
    rec = s.knots[L["profileOffset"]:]
    w, h = int(rec[0]), int(rec[1])
    return rec[8:8 + w * h * 3].reshape(h, w, 3)


@pytest.mark.parametrize("size,rle", [((5, 11), True), ((4, 6), True), ((3, 9), False)])
def test_radiance_hdr_maps(host, tmpdir, size, rle):
    """Radiance RGBE written by the format's definition (run-length coded scanlines, or flat; a
    scanline narrower than 8 texels is always flat): each texel is c * 2^(e - 136), and that is
    the map's colour within RGBE's 8-bit mantissas under the texel's shared exponent."""
    h, w = size
    rgb = (np.random.default_rng(w).random((h, w, 3)) ** 3 * 50).astype(np.float32)
    rgb[0, 0] = 0
    rgbe = ts.write_hdr(f"{tmpdir}/m{w}.hdr", rgb, rle=rle)
    got = _map_texels(host, tmpdir, f"hdr{w}", f"{tmpdir}/m{w}.hdr")
    e = rgbe[..., 3].astype(np.int64)
    want = np.where(e[..., None] > 0, rgbe[..., :3] * np.ldexp(np.float32(1), e - 136)[..., None], 0).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    # the shared exponent quantises every component to 1/128 of the texel's largest one
    assert np.all(np.abs(got - rgb) <= rgb.max(axis=2, keepdims=True) / 128 + 1e-30)


@pytest.mark.parametrize("compression", [0, 1, 2, 3])
@pytest.mark.parametrize("dtype,layout", [(np.float16, "RGB"), (np.float32, "RGBA"), (np.uint32, "RGB"),
                                          (np.float16, "Y")])
def test_openexr_maps(host, tmpdir, compression, dtype, layout):
    """Single-part scanline OpenEXR written by the file layout (channels in name order, NONE /
    RLE / ZIPS / ZIP with the byte predictor and interleave, 16-line ZIP blocks, a data window
    away from the origin, bottom-first chunk order for half the cases): the map's texels are the
    R, G, B channels (a lone channel as grey; UINT bits read as floats; a DECREASING_Y file upside
    down, as the reference's reader returns it), exactly."""
    rng = np.random.default_rng(compression * 10 + len(layout))
    h, w = 37, 13
    base = rng.random((h, w)) * 4
    base[5:20] = 1.5  # runs for RLE
    if dtype == np.uint32:
        chans = {c: (base * (k + 1) * 1000).astype(np.uint32) for k, c in enumerate(layout)}
    else:
        chans = {c: (base * (k + 1)).astype(dtype) for k, c in enumerate(layout)}
    name = f"e{compression}{np.dtype(dtype).char}{layout}"
    ts.write_exr(f"{tmpdir}/{name}.exr", chans, compression, window=(3, -2), decreasing=compression % 2 == 1)
    got = _map_texels(host, tmpdir, name, f"{tmpdir}/{name}.exr")
    # a UINT channel's bits are taken as a float, as the reference's LoadEXR returns them
    as_float = (lambda a: a.view(np.float32)) if dtype == np.uint32 else (lambda a: a.astype(np.float32))
    want = np.stack([as_float(chans["Y" if layout == "Y" else c]) for c in "RGB"], axis=2)
    if compression % 2 == 1:  # DECREASING_Y: the reference's LoadEXR returns it upside down
        want = want[::-1]
    np.testing.assert_array_equal(got, want)


def test_openexr_and_hdr_errors(host, tmpdir):
    chans = {c: np.ones((4, 4), np.float16) for c in "RGB"}
    ts.write_exr(f"{tmpdir}/piz.exr", chans, 4)
    with pytest.raises(RuntimeError, match="unsupported OpenEXR compression 4"):
        es.load(host, tmpdir, "piz", es.env_doc(f"{tmpdir}/piz.exr", objects=False))
    ts.write_exr(f"{tmpdir}/uv.exr", {"U": np.ones((4, 4), np.float16), "V": np.ones((4, 4), np.float16)})
    with pytest.raises(RuntimeError, match="without R, G and B"):
        es.load(host, tmpdir, "uv", es.env_doc(f"{tmpdir}/uv.exr", objects=False))
    data = bytearray(open(f"{tmpdir}/uv.exr", "rb").read())
    data[5] |= 0x02  # the tiled flag (bit 9 of the version field)
    open(f"{tmpdir}/tiled.exr", "wb").write(bytes(data))
    with pytest.raises(RuntimeError, match="not tiled"):
        es.load(host, tmpdir, "tiled", es.env_doc(f"{tmpdir}/tiled.exr", objects=False))
    open(f"{tmpdir}/bad.hdr", "wb").write(b"#?RADIANCE\n\n+Y 4 +X 4\n" + bytes(64))
    with pytest.raises(RuntimeError, match="layout"):
        es.load(host, tmpdir, "badhdr", es.env_doc(f"{tmpdir}/bad.hdr", objects=False))


@pytest.mark.parametrize("tag", ["hdr", "exr"])
def test_float_maps_equal_the_reference_decoders(host, tmpdir, tag):
    """The reference reads .hdr with its vendored stb_image float loader and .exr with its vendored
    tinyexr LoadEXR (src/core/image.cpp:79-89, src/core/exr_read.cpp:19-37). The golden vectors
    hold those decoders' texels for Radiance files (run-length coded, flat, narrow) and scanline
    OpenEXR files (NONE / RLE / ZIPS / ZIP, HALF / FLOAT / UINT, RGB / RGBA / one grey channel,
    an offset data window, both line orders): the maps' texels equal them bit for bit."""
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_components.npz"))
    ends, shapes = G[f"{tag}_file_ends"], G[f"{tag}_shapes"]
    starts = np.concatenate([[0], ends[:-1]])
    t0 = 0
    for k, (a, b) in enumerate(zip(starts, ends)):
        path = f"{tmpdir}/golden{k}.{tag}"
        with open(path, "wb") as f:
            f.write(G[f"{tag}_file_bytes"][a:b].tobytes())
        h, w = shapes[k]
        want = G[f"{tag}_texels"][t0:t0 + h * w * 3].reshape(h, w, 3)
        t0 += h * w * 3
        np.testing.assert_array_equal(_map_texels(host, tmpdir, f"g{tag}{k}", path), want, err_msg=f"{tag} file {k}")
