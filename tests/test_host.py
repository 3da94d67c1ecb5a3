"""Host library: scene generators, named optical data, RGB uplift and film resolve. CPU only."""
import os

import numpy as np
import pytest

import prender_amd as pa

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_components.npz"))


@pytest.fixture(scope="module")
def c2_small(host):
    return host.generate(2, 96, 54, 0.05)


def test_generators_are_deterministic(host):
    a, b = host.generate(1, 32, 32), host.generate(1, 32, 32)
    for f in ("prims", "materials", "spectra", "knots", "lights"):
        assert np.array_equal(getattr(a, f).view(np.uint8), getattr(b, f).view(np.uint8)), f


def test_configuration_1_shape(c1):
    assert (c1.width, c1.height) == (256, 256)
    assert c1.defaults["spp"] == 64 and c1.defaults["maxDepth"] == 8 and c1.defaults["rrStartDepth"] == 3
    assert len(c1.lights) == 2 and np.all(c1.lights["kind"] == pa.LIGHT_AREA)
    # every area light's primitive points back at its light
    assert np.array_equal(c1.prims["light"][c1.lights["primitive"]], np.arange(len(c1.lights)))


def test_configuration_2_full_shape(host):
    s = host.generate(2, 64, 36)
    tris = (s.prims["type"] == pa.PRIM_TRIANGLE).sum()
    assert tris == 999 * 499 * 2 + 200 * 5120 + 64 * 2
    assert len(s.lights) == 128
    assert s.defaults["spp"] == 1024 and s.defaults["maxDepth"] == 16
    mt = s.materials["type"]
    sphere_mats = mt[np.unique(s.prims["material"][~np.isin(np.arange(len(s.prims)), s.lights["primitive"])])]
    assert {pa.MAT_DIFFUSE, pa.MAT_CONDUCTOR, pa.MAT_DIELECTRIC} <= set(sphere_mats.tolist())


def test_named_metals_and_glass_match_reference_data(host, c2_small):
    """The generator's conductor n/k and BK7 IOR equal the reference's named data at every sampled wavelength."""
    s, lam = c2_small, G["spec_lambda"]

    def values(spec):
        return np.array([host.spectrum_value(s.spectra[spec], s.knots, float(l)) for l in lam], np.float32)

    cond = s.materials[s.materials["type"] == pa.MAT_CONDUCTOR]
    assert len(cond) > 0
    for m in cond:
        eta, k = values(m["eta"]), values(m["k"])
        d = [max(np.abs(eta - G["metal_values"][i, 0]).max(), np.abs(k - G["metal_values"][i, 1]).max())
             for i in range(4)]
        assert min(d) < 1e-5, d
    diel = s.materials[s.materials["type"] == pa.MAT_DIELECTRIC]
    assert len(diel) > 0
    for m in diel:
        np.testing.assert_allclose(values(m["eta"]), G["glass_values"][0], rtol=1e-6)
        assert m["dispersive"] == 1


def test_blackbody_lights_have_unit_luminance_shape(host, c2_small):
    s = c2_small
    sp = s.spectra[s.lights["spectrum"]]
    assert np.all(sp["kind"] == pa.SPEC_BLACKBODY)
    T = sp["p"][:, 0]
    assert T.min() >= 2700 and T.max() <= 9000
    # the reference's BlackbodySpectrum normalised to Y = 1, at the temperatures of the goldens
    for t_idx, T0 in enumerate(G["bb_temps"]):
        row = np.zeros(1, pa.SPECTRUM_DTYPE)
        row["kind"], row["scale"], row["p"][0, 0] = pa.SPEC_BLACKBODY, 1.0, T0
        lam_peak = 2.8977721e-3 / T0
        c, h, kb = 299792458.0, 6.62606957e-34, 1.3806488e-23
        row["p"][0, 1] = 1 / ((2 * h * c * c) / (lam_peak ** 5 * (np.exp((h * c) / (lam_peak * kb * T0)) - 1)))
        vals = np.array([host.spectrum_value(row[0], s.knots, float(l)) for l in G["spec_lambda"]])
        assert np.all(vals > 0)
        ratio = G["bb_unit_y"][t_idx] / vals
        np.testing.assert_allclose(ratio, ratio[0], rtol=2e-5)  # same shape; the scale is the unit-Y factor


def _lab(rgb, to_xyz, white):
    """CIELAB of linear-sRGB rows under the fit's matrix and white (the fit's error measure)."""
    xyz = rgb @ to_xyz.T / white
    d = 6 / 29
    f = np.where(xyz > d ** 3, np.cbrt(xyz), xyz / (3 * d * d) + 4 / 29)
    return np.stack([116 * f[:, 1] - 16, 500 * (f[:, 0] - f[:, 1]), 200 * (f[:, 1] - f[:, 2])], 1)


def _table(host):
    rec = host.rgb_table_record()
    n = int(rec[0])
    assert n == 64 and rec.size == 1 + n + 3 * n ** 3 * 3
    return n, rec[1:1 + n], rec[1 + n:].reshape(3, n, n, n, 3)


def test_rgb_fit_inputs_equal_the_reference(host):
    """The objective the RGB-to-spectrum table solves: per fine wavelength the spectrum-to-sRGB
    weights (D65, the spectral white adaptation, the sRGB matrix, trapezoid weights, 1 / CIE Y
    integral), the sRGB-to-XYZ matrix and its white. The goldens are computed from the reference's
    own compiled colour and spectrum code (src/core/rgb2spec.cpp:39-74): bit for bit."""
    w, m, white = host.rgb_fit_tables()
    assert np.array_equal(w, G["fit_rgbw"]) and np.array_equal(m, G["fit_to_xyz"]) and np.array_equal(white, G["fit_white"])


def test_rgb_table_nodes_solve_the_reference_objective(host):
    """Each table node's coefficients (in nm, as stored) reproduce the node's colour under the
    reference's fit inputs: CIELAB residual below 1e-3 for more than 99.9 % of the nodes from the
    brightness where the table's walk starts (z node res / 5) upwards; the others, above 2e-3, are
    the saturated edges (another component at most 2 / 63 of the maximum) and the corner by white,
    which a sigmoid reaches only in the limit. Three equations in three unknowns: where the residual
    vanishes the coefficients are the objective's root, the one the reference's own walk (the same
    start, the same steps) converges to."""
    n, zn, co = _table(host)
    w, to_xyz, white = G["fit_rgbw"], G["fit_to_xyz"], G["fit_white"]
    lam = 360.0 + 2.5 * np.arange(w.shape[1])
    rng = np.random.default_rng(7)
    l, k, j, i = rng.integers(3, size=40000), rng.integers(n // 5, n, 40000), rng.integers(n, size=40000), \
        rng.integers(n, size=40000)
    c = co[l, k, j, i].astype(np.float64)
    v = (c[:, :1] * lam + c[:, 1:2]) * lam + c[:, 2:3]
    rgb_fit = (0.5 * v / np.sqrt(1 + v * v) + 0.5) @ w.T
    z = zn[k].astype(np.float64)
    target = np.zeros((len(z), 3))
    rows = np.arange(len(z))
    target[rows, l], target[rows, (l + 1) % 3], target[rows, (l + 2) % 3] = z, i / (n - 1) * z, j / (n - 1) * z
    r = np.linalg.norm(_lab(target, to_xyz, white) - _lab(rgb_fit, to_xyz, white), axis=1)
    print(f"table nodes from z node {n // 5}: {np.mean(r < 1e-3):.5f} with CIELAB residual < 1e-3 "
          f"(median {np.median(r):.2e})", flush=True)
    assert np.mean(r < 1e-3) > 0.999
    lo = np.minimum(i, j)
    assert np.all(((lo <= 2) | ((k >= 54) & (lo >= 58)))[r >= 2e-3])


def test_rgb_lookup_at_node_colours_returns_the_nodes(host):
    """RGBToSpectrumTable::operator() at a node's own colour (z exactly a z node, the other two at
    i / (res - 1) and j / (res - 1) of it) interpolates back to that node: the maximum component
    picks the table, x and y index its columns in that order, FindInterval the z node."""
    n, zn, co = _table(host)
    rng = np.random.default_rng(8)
    lam = np.linspace(360, 830, 48)
    for _ in range(300):
        l, k, j, i = int(rng.integers(3)), int(rng.integers(1, n)), int(rng.integers(n - 1)), int(rng.integers(n - 1))
        z = zn[k]
        rgb = np.zeros(3, np.float32)
        rgb[l], rgb[(l + 1) % 3], rgb[(l + 2) % 3] = z, np.float32(i / (n - 1)) * z, np.float32(j / (n - 1)) * z
        c = host.fit_sigmoid(rgb).astype(np.float64)
        ref = co[l, k, j, i].astype(np.float64)
        sig = lambda q: 0.5 * ((q[0] * lam + q[1]) * lam + q[2]) / np.sqrt(1 + ((q[0] * lam + q[1]) * lam + q[2]) ** 2) + 0.5  # noqa: E731
        np.testing.assert_allclose(sig(c), sig(ref), atol=2e-4, err_msg=f"node {(l, k, j, i)}")


def test_sigmoid_fit_grey_and_round_trip(host):
    c = host.fit_sigmoid([0.5, 0.5, 0.5])
    row = np.zeros(1, pa.SPECTRUM_DTYPE)
    row["kind"], row["scale"] = pa.SPEC_SIGMOID, 1.0
    row["p"][0, :3] = c
    for l in (400.0, 550.0, 700.0):
        assert abs(host.spectrum_value(row[0], np.zeros(0, np.float32), l) - 0.5) < 1e-3
    # saturated red reflects long wavelengths and absorbs short ones
    c = host.fit_sigmoid([0.8, 0.1, 0.1])
    row["p"][0, :3] = c
    red = host.spectrum_value(row[0], np.zeros(0, np.float32), 650.0)
    blue = host.spectrum_value(row[0], np.zeros(0, np.float32), 450.0)
    assert red > 0.6 and blue < 0.2


def test_resolve_srgb_of_d65_white_is_white(host):
    xyz = np.array([[0.95047, 1.0, 1.08883]], np.float64)
    rgb = host.resolve_srgb(xyz, 1.0)
    np.testing.assert_allclose(rgb[0], 1.0, atol=2e-3)


def test_unavailable_configurations_fail_loudly(host):
    for cfg in (0, 7):
        with pytest.raises(RuntimeError):
            host.generate(cfg, 16, 16)


def test_configuration_5_gems_are_closed_convex_solids(host):
    s = host.generate(5, 64, 36)
    assert s.camera.lensRadius > 0 and s.defaults["spp"] == 4096
    assert {pa.LIGHT_AREA, pa.LIGHT_ENV} <= set(s.lights["kind"].tolist())
    diel = np.flatnonzero(s.materials["type"] == pa.MAT_DIELECTRIC)
    assert len(diel) == 1 and s.spectra[s.materials["eta"][diel[0]]]["kind"] == pa.SPEC_SELLMEIER
    tris = s.prims[s.prims["material"] == diel[0]]
    per_gem = len(tris) // 150
    assert per_gem * 150 == len(tris)
    P = np.stack([tris["p0"], tris["p1"], tris["p2"]], 1).astype(np.float64)
    # every directed edge has exactly one reverse twin (closed, consistently wound)
    edges = {}
    for t in P:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = (tuple(t[a]), tuple(t[b]))
            edges[key] = edges.get(key, 0) + 1
    assert all(c == 1 and edges.get((k[1], k[0]), 0) == 1 for k, c in edges.items())
    # each gem is convex with outward normals: all its vertices lie behind every facet plane
    for g in range(0, 150, 37):
        G = P[g * per_gem:(g + 1) * per_gem]
        n = np.cross(G[:, 1] - G[:, 0], G[:, 2] - G[:, 0])
        verts = G.reshape(-1, 3)
        dist = np.einsum("fk,fvk->fv", n, verts[None, :, :] - G[:, 0, None, :])
        assert dist.max() <= 1e-4 * np.abs(n).max()


def test_configuration_6_coated_and_sheen_materials(host):
    s = host.generate(6, 64, 36, 0.25)
    m = s.materials
    coated = np.isin(m["type"], [pa.MAT_COATED_DIFFUSE, pa.MAT_COATED_CONDUCTOR])
    assert coated.sum() >= 20 and (m["type"] == pa.MAT_COATED_CONDUCTOR).sum() >= 6
    assert (m["sheenWeight"] > 0).sum() >= 6 and np.all(m["sheenColor"][m["sheenWeight"] > 0] >= 0)
    # a scattering slab, a rough coat and a dispersive (Sellmeier) coat are all present
    assert np.any(m["coatAlbedo"][coated] >= 0) and np.any(m["coatAlpha"][coated] > 0)
    kinds = s.spectra["kind"][m["coatEta"][coated]]
    assert pa.SPEC_SELLMEIER in kinds.tolist() and np.any(m["coatDispersive"][coated] == 1)
    assert np.all(m["thickness"][coated] > 0) and np.all(m["coatMaxDepth"][coated] == 64)


def test_checkpoint_round_trip_and_errors(host, tmp_path):
    """The reference checkpoint format (src/integrators/integrator.cpp:16-90) and its three error
    messages; the film and the pass count survive a write/read round trip exactly."""
    s = host.generate(1, 24, 16)
    rng = np.random.default_rng(3)
    film, sq = rng.random((16, 24, 3)), rng.random((16, 24, 3))
    path = tmp_path / "render.ckpt"
    host.checkpoint_write(path, s, 8, 1, film, sq, 37)
    raw = path.read_bytes()
    assert raw[:8] == b"PRCKPT02"
    assert int.from_bytes(raw[8:12], "little") == 0x100  # path integrator on a GPU device
    assert int.from_bytes(raw[44:52], "little") == 24 * 16 * 3  # 44-byte header, then n
    f2, q2, passes = host.checkpoint_read(path, s, 8, 1)
    assert passes == 37 and np.array_equal(f2, film) and np.array_equal(q2, sq)
    # a reader that does not want the variance film still gets the sums and passes
    f3, _, p3 = host.checkpoint_read(path, s, 8, 1, variance=False)
    assert p3 == 37 and np.array_equal(f3, film)
    with pytest.raises(RuntimeError, match="different scene"):
        host.checkpoint_read(path, s, 9, 1)
    with pytest.raises(RuntimeError, match="different scene"):
        host.checkpoint_read(path, s, 8, 2)
    other = host.generate(1, 24, 16, 0.5)
    assert host.fingerprint(other) != host.fingerprint(s)
    with pytest.raises(RuntimeError, match="different scene"):
        host.checkpoint_read(path, other, 8, 1)
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"not a checkpoint at all, just bytes" * 3)
    with pytest.raises(RuntimeError, match="not a prender checkpoint"):
        host.checkpoint_read(bad, s, 8, 1)
    cut = tmp_path / "cut.ckpt"
    cut.write_bytes(raw[:len(raw) // 2])
    with pytest.raises(RuntimeError, match="truncated or corrupt"):
        host.checkpoint_read(cut, s, 8, 1)
    assert not (tmp_path / "render.ckpt.tmp").exists()


def test_preetham_sky_matches_reference_maps(host):
    """The Preetham daylight map of sun_sky lights against the reference's MakePreethamSky
    (src/lights/sky.cpp:146-208) at 48 x 24 texels: sun high, low, just below the horizon, at the
    zenith, turbidity clamped from both sides. Same float operations in the same order, so the
    maps agree to a few ulps (the ground term sums the sky's irradiance in double)."""
    for case, ref in zip(G["sky_cases"], G["sky_maps"]):
        sky = host.preetham_sky(case[:3], float(case[3]), float(case[4]), 48, 24)
        np.testing.assert_allclose(sky, ref, rtol=1e-5, atol=1e-6, err_msg=str(case))
        # Independent of the reference: the lower half is one grey ground value, the sky is
        # non-negative, and a sun below -0.1 in y leaves no daylight above the horizon.
        ground = sky[12:]
        assert np.all(ground == ground[0, 0, 0]) and np.all(sky >= 0)
        if case[1] / np.linalg.norm(case[:3]) <= -0.1:
            assert not sky[:12].any()
    with pytest.raises(ValueError):
        host.preetham_sky((0, 1, 0), 3.0, 0.2, 0, 8)


def test_envmap_sampling_matches_reference(host):
    """Image-environment sampling tables (luminance x sin(theta) + 1e-6 weights, per-row
    conditional and marginal CDFs in double) against the reference's Distribution2D: the sampled
    (u, v), its pdf and the density are bit-identical for 2000 random pairs and the corners. The
    density also matches the reference EnvironmentLight's own PDF_Li, which builds its own table,
    after the change of variables pdf / (2 pi^2 sin theta), away from the poles (where the
    reference's round trip through a direction amplifies rounding)."""
    rgb, u, ref = G["sky_maps"][0], G["env_u"], G["env_ref"]
    got, avg_y = host.envmap_sample(rgb, u)
    assert np.array_equal(got, ref[:, :4])
    sin_t = np.sin(got[:, 1] * np.pi)
    away = sin_t > 0.05
    pdf_li = got[away, 3] / (2 * np.pi ** 2 * sin_t[away])
    np.testing.assert_allclose(pdf_li, ref[away, 4], rtol=1e-4)
    # Independent of the reference: the density of a sample is its cell's weight over the mean
    # weight (so it integrates to one over the unit square), the sampled pdf is that density, and
    # the mean luminance is the sin-weighted mean.
    h, w = rgb.shape[:2]
    sin_rows = np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    lum = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    f = np.maximum(lum, 0) * sin_rows + 1e-6
    iu = np.clip((got[:, 0] * w).astype(int), 0, w - 1)
    iv = np.clip((got[:, 1] * h).astype(int), 0, h - 1)
    np.testing.assert_allclose(got[:, 3], f[iv, iu] / f.mean(), rtol=1e-5)
    np.testing.assert_allclose(got[:, 2], got[:, 3], rtol=1e-6)
    assert avg_y == pytest.approx((lum * sin_rows).sum() / (sin_rows.sum() * w), rel=1e-5)
    with pytest.raises(ValueError):
        host.envmap_sample(rgb[:0], u)
