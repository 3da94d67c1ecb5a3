"""The oracle against what the mathematics fixes: normalisation integrals, closed forms, special
cases that reduce to textbook formulas, and brute force. These pin the parts the reference's
golden vectors leave open (media, energy conservation, BVH against exhaustive search). CPU only.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
import prender_amd as pa
import raygen

L = orc.lib()
_alive = []


def p(a):
    a = np.ascontiguousarray(a)
    _alive.append(a)
    return a.ctypes.data


@pytest.fixture(autouse=True)
def _release_buffers():
    yield
    _alive.clear()


def f32(*v):
    return np.ascontiguousarray(np.array(v, np.float32).ravel())


def uniform_sphere(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3)).astype(np.float32)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def bxdf(desc, op, wo, wi, uc=0.5, u=(0.5, 0.5)):
    out = np.zeros(11, np.float32)
    L.or_bxdf(p(f32(*desc)), op, p(f32(*wo)), p(f32(*wi)), float(uc), p(f32(*u)), p(out))
    return out


def test_sample_visible_pdf_normalised_and_inverse():
    lam = np.linspace(360, 830, 20001)
    pdf = 0.0039398042 / np.cosh(0.0072 * (lam - 538)) ** 2  # SampleVisible density (spectrum.cpp:11-19)
    assert abs(np.trapezoid(pdf, lam) - 1) < 2e-3
    for u in (0.1, 0.37, 0.5, 0.9):
        l4, p4 = np.zeros(4, np.float32), np.zeros(4, np.float32)
        L.or_sample_visible(u, p(l4), p(p4))
        np.testing.assert_allclose(p4, 0.0039398042 / np.cosh(0.0072 * (l4 - 538)) ** 2, rtol=1e-5)
        # the four hero wavelengths are u + k/4 (mod 1) mapped through the inverse CDF
        cdf = lambda x: np.interp(x, lam, np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(lam))]))
        diff = (cdf(l4) - np.array([u + k / 4 for k in range(4)])) % 1  # circular: u = 1 is kept as 1
        assert np.all(np.minimum(diff, 1 - diff) < 2e-3), diff


def test_cie_y_integral_and_white():
    lam = np.arange(360, 831, 0.25, dtype=np.float32)
    y = np.zeros(len(lam))
    for i, l in enumerate(lam):
        out = np.zeros(3, np.float32)
        L.or_cie(float(l), p(out))
        y[i] = out[1]
    np.testing.assert_allclose(np.trapezoid(y, lam), L.or_cie_y_integral(), rtol=1e-4)
    # Unit radiance at all wavelengths has Y = 1 in expectation over the hero-wavelength sampling.
    acc = 0.0
    us = (np.arange(4096) + 0.5) / 4096
    for u in us:
        l4, p4, xyz = np.zeros(4, np.float32), np.zeros(4, np.float32), np.zeros(3, np.float32)
        L.or_sample_visible(float(u), p(l4), p(p4))
        L.or_to_xyz(p(np.ones(4, np.float32)), p(l4), p(p4), p(xyz))
        acc += xyz[1]
    assert abs(acc / len(us) - 1) < 3e-3


def spec_row(kind, params, scale=1.0, offset=0, count=0):
    r = np.zeros(1, pa.SPECTRUM_DTYPE)
    r["kind"], r["scale"], r["offset"], r["count"] = kind, scale, offset, count
    r["p"][0, :len(params)] = params
    return r


def test_sellmeier_and_sigmoid_known_values():
    bk7 = spec_row(pa.SPEC_SELLMEIER, (1.03961212, 0.231792344, 1.01046945, 0.00600069867, 0.0200179144, 103.560653),
                   count=3)
    n_d = L.or_spectrum_eval(p(bk7), None, 0, 587.56)
    assert abs(n_d - 1.5168) < 1e-4  # Schott BK7 n_d
    grey = spec_row(pa.SPEC_SIGMOID, (0, 0, 0))
    assert L.or_spectrum_eval(p(grey), None, 0, 500.0) == 0.5
    sat = spec_row(pa.SPEC_SIGMOID, (0, 0, 1e4))
    assert L.or_spectrum_eval(p(sat), None, 0, 500.0) > 0.999


def test_blackbody_unit_luminance(host):
    """A blackbody normalised to unit luminance integrates to Y = 1 against the CIE fit."""
    for T in (2700.0, 6500.0):
        peak = np.float32(2.8977721e-3 / T)
        c, h, kb = 299792458.0, 6.62606957e-34, 1.3806488e-23
        planck = lambda lm: (2 * h * c * c) / (lm ** 5 * (np.exp((h * c) / (lm * kb * T)) - 1))
        row = spec_row(pa.SPEC_BLACKBODY, (T, 1 / planck(float(peak))))
        y = L.or_spectrum_y(p(row), None, 0)
        row["scale"] = 1 / y
        assert abs(L.or_spectrum_y(p(row), None, 0) - 1) < 1e-5
        assert abs(L.or_spectrum_eval(p(row), None, 0, float(peak) * 1e9) * y - 1) < 1e-4


def test_fresnel_special_cases():
    for eta in (1.33, 1.5, 2.4):
        assert abs(L.or_fr_dielectric(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
        # k = 0 conductor Fresnel is the dielectric one
        for c in (0.1, 0.5, 0.9):
            assert abs(L.or_fr_complex(c, eta, 0.0) - L.or_fr_dielectric(c, eta)) < 1e-5
    # total internal reflection from inside glass beyond the critical angle
    assert L.or_fr_dielectric(0.2, 1 / 1.5) == 1.0
    # entering from the back side is the same as eta -> 1/eta from the front
    assert abs(L.or_fr_dielectric(-0.8, 1.5) - L.or_fr_dielectric(0.8, 1 / 1.5)) < 1e-6


def test_hg_normalised_and_isotropic():
    mu = np.linspace(-1, 1, 40001)
    for g in (-0.7, 0.0, 0.3, 0.9):
        vals = np.array([L.or_hg(float(m), g) for m in mu[::10]])
        assert abs(2 * np.pi * np.trapezoid(vals, mu[::10]) - 1) < 2e-3, g
    assert abs(L.or_hg(0.3, 0.0) - 1 / (4 * np.pi)) < 1e-7


def test_bxdf_pdf_integral_equals_sampling_success_rate():
    """Monte Carlo over the sphere: each BxDF's pdf integrates to the probability that Sample_f
    returns a sample (directions that leave the valid hemisphere are rejected by both), and the
    pdf a sample reports is the pdf evaluated at its direction."""
    wi = uniform_sphere(40000, 3)
    wo = f32(0.3, -0.2, 0.932738)
    wo /= np.linalg.norm(wo)
    rng = np.random.default_rng(4)
    for desc in ([1, 0, 0, 1, .5, .5, .5, .5] + [0] * 8,
                 [2, 0.3, 0.3, 1, 0, 0, 0, 0, .2, .3, .4, .5, 3, 3.2, 3.4, 3.6],
                 [2, 0.1, 0.5, 1, 0, 0, 0, 0, .2, .3, .4, .5, 3, 3.2, 3.4, 3.6],
                 [3, 0.3, 0.3, 1.5] + [0] * 12):
        est = np.array([bxdf(desc, 1, wo, w)[0] for w in wi], np.float64) * 4 * np.pi
        mean, err = est.mean(), est.std() / np.sqrt(len(est))
        ok = 0
        n = 20000
        for k in range(n):
            s = bxdf(desc, 2, wo, wo, rng.random(), rng.random(2))
            if s[10]:
                ok += 1
                if k < 200:
                    np.testing.assert_allclose(s[7], bxdf(desc, 1, wo, s[4:7])[0], rtol=2e-4)
        rate = ok / n
        tol = 4 * np.hypot(err, np.sqrt(rate * (1 - rate) / n)) + 2e-3
        assert abs(mean - rate) < tol, (desc[:4], mean, rate, tol)
        assert rate > 0.85


def test_lambert_sampling_weight_is_albedo():
    R = (0.2, 0.4, 0.6, 0.8)
    rng = np.random.default_rng(5)
    for _ in range(50):
        u = rng.random(2)
        s = bxdf([1, 0, 0, 1, *R] + [0] * 8, 2, (0, 0, 1), (0, 0, 1), 0.5, u)
        assert s[10] == 1
        np.testing.assert_allclose(s[:4] * abs(s[6]) / s[7], R, rtol=1e-5)


def test_smooth_conductor_reflects_mirror_direction():
    wo = f32(0.6, 0.0, 0.8)
    s = bxdf([2, 0, 0, 1, 0, 0, 0, 0, .2, .3, .4, .5, 3, 3.2, 3.4, 3.6], 2, wo, wo)
    np.testing.assert_allclose(s[4:7], (-0.6, 0, 0.8), atol=1e-6)
    fr = [L.or_fr_complex(0.8, e, k) for e, k in zip((.2, .3, .4, .5), (3, 3.2, 3.4, 3.6))]
    np.testing.assert_allclose(s[:4] * 0.8 / s[7], fr, rtol=1e-5)


def test_bvh_matches_brute_force(c1):
    o = orc.OracleScene(c1)
    lo, hi = raygen.scene_bounds(c1)
    rays = np.concatenate([raygen.random_rays(lo, hi, 4000, 17), raygen.camera_rays(o, c1, 2000, 18)])
    hits = o.trace(rays)
    brute, ties = o.trace_brute(rays)
    keep = ties == 0
    assert np.array_equal(hits["prim"][keep], brute["prim"][keep])
    assert np.array_equal(hits["t"], brute["t"])
    occl = o.trace(rays, any_hit=True)
    assert np.array_equal(occl["prim"] >= 0, brute["prim"] >= 0)


def test_light_bvh_pmfs_are_the_sampling_probabilities(host):
    """The PMFs of all lights add up to the probability that the light-BVH descent returns a light:
    below 1 where a node's bounds give importance while both its children give none (the descent
    then stops with no light, reference light_bvh.cpp:179-214). Checked against a stratified sweep
    of u, whose error is at most (2 x leaves) / samples; every sampled light's pmf equals its PMF."""
    s = host.generate(2, 64, 36, 0.05)
    o = orc.OracleScene(s)
    rng = np.random.default_rng(9)
    lo, hi = raygen.scene_bounds(s)
    nu = 4096
    for _ in range(6):
        pt = rng.uniform(lo, hi).astype(np.float32)
        n = uniform_sphere(1, int(rng.integers(1 << 30)))[0]
        pm = np.array([o.light_pmf(pt, n, k) for k in range(len(s.lights))])
        assert pm.sum() <= 1 + 1e-5
        hits = 0
        for u in (np.arange(nu) + 0.5) / nu:
            li, pmf = o.light_sample(pt, n, float(u))
            if li >= 0:
                hits += 1
                assert abs(pmf - pm[li]) <= 1e-6 * max(1, pmf)
        assert abs(hits / nu - pm.sum()) <= 2 * len(s.lights) / nu
    pw = np.array([o.light_pmf((0, 0, 0), (0, 0, 0), k, bvh=False) for k in range(len(s.lights))])
    assert abs(pw.sum() - 1) < 1e-5


def medium_scene(kind, density=None, sigma_a=0.3, sigma_s=0.5, max_density=1.0):
    prims = np.zeros(1, pa.PRIM_DTYPE)
    prims["p0"], prims["p1"], prims["p2"] = (100, 100, 100), (101, 100, 100), (100, 101, 100)
    prims["mediumInside"] = prims["mediumOutside"] = -1
    prims["light"] = -1
    mats = np.zeros(1, pa.MATERIAL_DTYPE)
    mats["albedo"] = 0
    mats["albedoB"] = mats["eta"] = mats["k"] = mats["coatEta"] = mats["coatAlbedo"] = mats["sheenColor"] = -1
    specs = np.concatenate([spec_row(pa.SPEC_CONSTANT, (0.5,)), spec_row(pa.SPEC_CONSTANT, (sigma_a,)),
                            spec_row(pa.SPEC_CONSTANT, (sigma_s,))])
    med = np.zeros(1, pa.MEDIUM_DTYPE)
    med["kind"], med["sigmaA"], med["sigmaS"], med["scale"], med["g"] = kind, 1, 2, 1.0, 0.0
    grid = None
    if kind == pa.MEDIUM_GRID:
        grid = np.full(4 * 4 * 4, density, np.float32)
        med["nx"] = med["ny"] = med["nz"] = 4
        med["maxDensity"] = max_density
        med["boundsMin"], med["boundsMax"] = (0, 0, 0), (1, 1, 1)
        med["worldToMedium"] = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)
    cam = pa.CameraEx()
    cam.cameraToWorld[:] = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)
    cam.tanHalfFov, cam.width, cam.height, cam.filterType, cam.filterRadius = 0.5, 4, 4, pa.FILTER_BOX, 0.5
    return pa.Scene(prims, mats, specs, None, None, med, grid, cam)


def lambdas(u=0.3):
    l4, p4 = np.zeros(4, np.float32), np.zeros(4, np.float32)
    L.or_sample_visible(u, p(l4), p(p4))
    return l4, p4


def test_homogeneous_beer_lambert_and_distance_sampling():
    s = medium_scene(pa.MEDIUM_HOMOGENEOUS)
    o = orc.OracleScene(s)
    l4, p4 = lambdas()
    for d in (0.1, 1.0, 3.0):
        out = np.zeros(4, np.float32)
        L.or_medium_tr(o.h, 0, p(f32(0, 0, 0)), p(f32(1, 0, 0)), d, p(l4), p(p4), p(out))
        np.testing.assert_allclose(out, np.exp(-0.8 * d), rtol=1e-6)
    # P(scatter before t) = 1 - exp(-sigma_t t); a scattering event's weight is sigma_s / sigma_t.
    us = (np.arange(2000) + 0.5) / 2000
    scat, w = 0, []
    for u in us:
        out = np.zeros(10, np.float32)
        L.or_medium_sample(o.h, 0, p(f32(0, 0, 0)), p(f32(1, 0, 0)), 1.5, float(u), p(l4), p(p4), p(out))
        if out[0]:
            scat += 1
            w.append(out[2:6])
            assert out[1] < 1.5
    assert abs(scat / len(us) - (1 - np.exp(-0.8 * 1.5))) < 2e-3
    np.testing.assert_allclose(np.array(w), 0.5 / 0.8, rtol=1e-5)


@pytest.mark.parametrize("density", [1.0, 0.5])
def test_grid_ratio_tracking_unbiased(density):
    s = medium_scene(pa.MEDIUM_GRID, density=density)
    o = orc.OracleScene(s)
    l4, p4 = lambdas()
    rng = np.random.default_rng(4)
    n, T = 6000, []
    for _ in range(n):
        y, z = rng.uniform(0.05, 0.95, 2)
        out = np.zeros(4, np.float32)
        L.or_medium_tr(o.h, 0, p(f32(-1, y, z)), p(f32(1, 0, 0)), 5.0, p(l4), p(p4), p(out))
        T.append(out[0])
    T = np.array(T, np.float64)
    expect = np.exp(-0.8 * density * 1.0)  # unit path length inside the box
    assert abs(T.mean() - expect) < 4 * T.std() / np.sqrt(n) + 1e-3, (T.mean(), expect)


def test_grid_spectral_tracking_scatter_probability():
    s = medium_scene(pa.MEDIUM_GRID, density=0.5)
    o = orc.OracleScene(s)
    l4, p4 = lambdas()
    rng = np.random.default_rng(8)
    n, scat, absorbed = 6000, 0, 0
    for _ in range(n):
        y, z = rng.uniform(0.05, 0.95, 2)
        out = np.zeros(10, np.float32)
        L.or_medium_sample(o.h, 0, p(f32(-1, y, z)), p(f32(1, 0, 0)), 5.0, float(rng.random()), p(l4), p(p4), p(out))
        if out[0]:
            scat += 1
            np.testing.assert_allclose(out[2:6], 1.0, rtol=1e-5)  # grey medium: unit weight
        elif not np.any(out[2:6]):
            absorbed += 1
    interact = 1 - np.exp(-0.8 * 0.5)
    for got, want in ((scat / n, interact * 0.5 / 0.8), (absorbed / n, interact * 0.3 / 0.8)):
        assert abs(got - want) < 4 * np.sqrt(want * (1 - want) / n), (got, want)


# ---- coated (layered) and sheen closures -----------------------------------------------------------
closure_desc = orc.closure_desc


def sample_weights(desc, wo, n, seed):
    """Sample_f weights f |cos| / pdf (zero for failed samples): their mean is the directional albedo.
    A walk through many near-specular rough bounces can overflow f and pdf in FP32 (as in the
    reference); such samples, about 1 in 10^4, are dropped as the film drops non-finite values."""
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 4))
    dropped = 0
    for k in range(n):
        s = orc.closure(desc, 2, wo, uc=rng.random(), u=rng.random(2))
        if s[10]:
            v = s[:4].astype(np.float64) * abs(s[6]) / s[7]
            if np.all(np.isfinite(v)):
                w[k] = v
            else:
                dropped += 1
    assert dropped <= max(2, n // 2000), dropped
    return w


def test_coated_white_diffuse_conserves_energy():
    """The slab between coat and base has unit extinction, so it absorbs exp(-thickness / |cos|)
    per crossing. A white Lambertian base under a clear smooth coat over a vanishingly thin slab
    absorbs nothing: light leaves after enough internal bounces and the albedo is 1 (radiance
    scaling 1/eta^2 in and eta^2 out cancel). The default thickness 0.01 loses a few percent."""
    wo = f32(0.3, -0.2, 0.932738)
    wo /= np.linalg.norm(wo)
    w = sample_weights(closure_desc(5, R=(1, 1, 1, 1), eta=1.5, thickness=1e-7), wo, 20000, 11)
    mean, err = w.mean(axis=0), w.std(axis=0) / np.sqrt(len(w))
    assert np.all(np.abs(mean - 1) < 4 * err + 0.005), (mean, err)
    thick = sample_weights(closure_desc(5, R=(1, 1, 1, 1), eta=1.5, thickness=0.01), wo, 8000, 12).mean()
    assert 0.85 < thick < 0.97
    # a grey base absorbs on every internal bounce: strictly less than its own albedo comes back
    w2 = sample_weights(closure_desc(5, R=(0.5, 0.5, 0.5, 0.5), eta=1.5, thickness=1e-7), wo, 8000, 12)
    assert w2.mean() < 0.5


@pytest.mark.parametrize("kind,kw", [
    (5, dict(R=(0.8, 0.4, 0.2, 0.6), coat_alpha=0.09, eta=1.33)),
    (5, dict(R=(0.8, 0.4, 0.2, 0.6), coat_alpha=0.25, thickness=0.3, albedo=(0.9, 0.5, 0.3, 0.7), g=0.4)),
    (6, dict(alpha=(0.1, 0.1), coat_alpha=0.04)),
])
def test_layered_f_integrates_to_sampled_albedo(kind, kw):
    """The stochastic f and the stochastic sampler estimate the same directional albedo:
    4 pi E_uniform[f |cos|] = E_sample[f |cos| / pdf] (rough coats: no delta lobes)."""
    desc = closure_desc(kind, **kw)
    wo = f32(-0.4, 0.1, 0.91104)
    wo /= np.linalg.norm(wo)
    wi = uniform_sphere(30000, 13)
    fv = np.array([orc.closure(desc, 0, wo, w)[:4] * abs(w[2]) for w in wi]) * 4 * np.pi
    ws = sample_weights(desc, wo, 20000, 14)
    a, ea = fv.mean(axis=0), fv.std(axis=0) / np.sqrt(len(fv))
    b, eb = ws.mean(axis=0), ws.std(axis=0) / np.sqrt(len(ws))
    assert np.all(np.abs(a - b) < 4 * np.hypot(ea, eb) + 0.01), (a, b, ea, eb)
    assert b.max() > 0.05


def concentric_cosine(u):
    """Cosine-weighted hemisphere directions by the concentric disk map (Shirley-Chiu)."""
    x, y = 2 * u[..., 0] - 1, 2 * u[..., 1] - 1
    big = np.abs(x) > np.abs(y)
    safe_x = np.where(x == 0, 1, x)
    safe_y = np.where(y == 0, 1, y)
    r = np.where(big, x, y)
    th = np.where(big, np.pi / 4 * (y / safe_x), np.pi / 2 - np.pi / 4 * (x / safe_y))
    dx, dy = r * np.cos(th), r * np.sin(th)
    z = np.sqrt(np.maximum(0, 1 - dx * dx - dy * dy))
    return np.stack([dx, dy, z], -1)


def test_charlie_sheen_lobe_normalised_and_tabulated():
    """On a black base the sheen closure is the Charlie lobe alone, f = D(wh) / (4 (ci + co - ci co)).
    (1) D is a normalised microfacet distribution: with wo = wi = wh, D = 4 f (2c - c^2) and
    the hemisphere integral of D(wh) cos(wh) is 1. (2) The albedo table's node (alpha_i, mu_j) is
    the 32 x 32 stratified cosine-weighted quadrature of f |cos| pi. (3) Where the lobe is broad
    that quadrature agrees with uniform Monte Carlo; at small alpha or grazing wo the peaked
    integrand makes the 32 x 32 rule read several percent off, as the reference's table does."""
    table = orc.sheen_table().reshape(16, 32)
    wh = uniform_sphere(40000, 15)
    wh[:, 2] = np.abs(wh[:, 2])
    wh = wh[wh[:, 2] > 1e-3]
    for alpha in (0.1, 0.3, 0.7, 1.0):
        desc = closure_desc(1, R=(0, 0, 0, 0), sheen=1.0, sheen_alpha=alpha)
        c = wh[:, 2].astype(np.float64)
        D = np.array([orc.closure(desc, 0, w, w)[0] for w in wh]) * 4 * (2 * c - c * c)
        est = 2 * np.pi * np.mean(D * c)
        err = 2 * np.pi * np.std(D * c) / np.sqrt(len(c))
        assert abs(est - 1) < 4 * err + 0.01, (alpha, est, err)
    g = (np.arange(32) + 0.5) / 32
    wi = concentric_cosine(np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)).astype(np.float32)
    for ia, im in ((2, 31), (7, 16), (15, 5), (0, 24), (11, 0)):
        alpha = float(np.float32(0.05 + 0.95 * ia / 15))
        mu = max(1e-3, im / 31)
        wo = f32(np.sqrt(1 - mu * mu), 0, mu)
        desc = closure_desc(1, R=(0, 0, 0, 0), sheen=1.0, sheen_alpha=alpha)
        quad = np.mean([orc.closure(desc, 0, wo, w)[0] for w in wi]) * np.pi
        np.testing.assert_allclose(table[ia, im], quad, rtol=2e-5, err_msg=f"node {ia} {im}")
        if ia >= 7 and im >= 5:  # broad lobe, wo away from grazing
            u = uniform_sphere(40000, 16 + ia)
            fv = np.array([orc.closure(desc, 0, wo, w)[0] * abs(w[2]) for w in u]) * 4 * np.pi
            est, err = fv.mean(), fv.std() / np.sqrt(len(fv))
            assert abs(est - table[ia, im]) < 4 * err + 0.02 * table[ia, im], (ia, im, est, table[ia, im], err)


def test_transport_mode_scales_dielectric_transmission():
    """Radiance transport divides refracted radiance by eta'^2; importance transport does not."""
    desc = closure_desc(3, alpha=(0.3, 0.3), eta=1.5)
    wo, wi = f32(0.2, 0.1, 0.974679), f32(-0.1, 0.05, -0.993734)
    fr = orc.closure(desc, 0, wo, wi, mode=0)[0]
    fi = orc.closure(desc, 0, wo, wi, mode=1)[0]
    assert fi > 0 and abs(fr / fi - 1 / 1.5 ** 2) < 1e-6
    # reflection-only and transmission-only sampling never return the other kind
    rng = np.random.default_rng(16)
    for rt, kind in ((1, 1), (2, 2)):
        for _ in range(200):
            s = orc.closure(desc, 2, wo, uc=rng.random(), u=rng.random(2), rt=rt)
            if s[10]:
                assert int(s[8]) & 3 == kind
