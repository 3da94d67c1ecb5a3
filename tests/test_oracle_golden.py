"""The oracle against golden vectors recorded from the reference's own compiled code.

tests/golden/ref_components.npz holds inputs and outputs of the reference's RNG stream
(src/integrators/integrator.h:29-38), SampledWavelengths::SampleVisible and the CIE fits
(src/core/spectrum.*), ToXYZ, the analytic spectra and named metal/glass data, the BxDFs
(src/materials/bxdf.*), Fresnel terms, Henyey-Greenstein, the pixel filters (src/core/film.*),
image textures and shading perturbation (src/materials/texture.cpp, material.cpp), and, on the
configuration-1 scene, the reference BVH (src/accel/bvh.cpp), light BVH
(src/lights/light_bvh.*), area-light sampling and the perspective camera. They were written by
tests/golden/make_goldens.py from oracle/_ref/libprref.so (the reference sources compiled as
they are). CPU only.
"""
import os

import numpy as np
import pytest

import oracle as orc
import prender_amd as pa

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_components.npz"))
L = orc.lib()


_alive = []


def p(a):
    """Address of a contiguous copy of `a`, kept alive until the next test starts."""
    a = np.ascontiguousarray(a)
    _alive.append(a)
    return a.ctypes.data


@pytest.fixture(autouse=True)
def _release_buffers():
    yield
    _alive.clear()


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def test_rng_stream_bit_exact():
    for (px, si, seed), ref in zip(G["rng_streams"], G["rng_values"]):
        out = np.zeros(16, np.float32)
        L.or_rng_stream(int(px), int(si), int(seed), 16, p(out))
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (px, si, seed)


def sample_visible(u):
    lam, pdf = np.zeros(4, np.float32), np.zeros(4, np.float32)
    L.or_sample_visible(float(u), p(lam), p(pdf))
    return lam, pdf


def test_sample_visible_and_cie():
    for u, rl, rp in zip(G["vis_u"], G["vis_lambda"], G["vis_pdf"]):
        lam, pdf = sample_visible(u)
        np.testing.assert_array_equal(lam, rl)
        np.testing.assert_array_equal(pdf, rp)
    for lam, ref in zip(G["cie_lambda"], G["cie_xyz"]):
        out = np.zeros(3, np.float32)
        L.or_cie(float(lam), p(out))
        np.testing.assert_array_equal(out, ref)
    assert np.float32(L.or_cie_y_integral()) == G["cie_y_integral"]


def test_to_xyz_including_terminated_secondaries():
    for u, Lv, term, ref in zip(G["xyz_u"], G["xyz_L"], G["xyz_term"], G["xyz_out"]):
        lam, pdf = sample_visible(u)
        if term:  # SampledWavelengths::TerminateSecondary (spectrum.h:111-115): secondaries' pdf 0, hero pdf / 4
            pdf = pdf.copy()
            pdf[1:] = 0
            pdf[0] /= 4
        out = np.zeros(3, np.float32)
        L.or_to_xyz(p(f32(Lv)), p(lam), p(pdf), p(out))
        np.testing.assert_allclose(out, ref, rtol=2e-6, atol=1e-12)


def test_analytic_spectra():
    knots = G["spec_knots"]
    for i, row in enumerate(G["spec_rows"]):
        r = np.ascontiguousarray(np.array([row], dtype=pa.SPECTRUM_DTYPE))
        vals = np.array([L.or_spectrum_eval(p(r), p(knots), len(knots), float(lam)) for lam in G["spec_lambda"]],
                        np.float32)
        np.testing.assert_allclose(vals, G["spec_values"][i], rtol=2e-6, atol=1e-12, err_msg=f"spectrum {i}")
        y = L.or_spectrum_y(p(r), p(knots), len(knots))
        np.testing.assert_allclose(y, G["spec_y"][i], rtol=1e-5, err_msg=f"spectrum {i} Y")


def test_bxdf_eval_pdf_sample():
    D, WO, WI, UC, U = G["bxdf_desc"], G["bxdf_wo"], G["bxdf_wi"], G["bxdf_uc"], G["bxdf_u"]
    worst = 0.0
    for i in range(len(D)):
        d = f32(D[i])
        for j in range(WO.shape[1]):
            out = np.zeros(11, np.float32)
            wo, wi, u = f32(WO[i, j]), f32(WI[i, j]), f32(U[i, j])
            L.or_bxdf(p(d), 0, p(wo), p(wi), float(UC[i, j]), p(u), p(out))
            np.testing.assert_allclose(out[:4], G["bxdf_f"][i, j], rtol=1e-5, atol=1e-7, err_msg=f"f {d[:4]}")
            L.or_bxdf(p(d), 1, p(wo), p(wi), float(UC[i, j]), p(u), p(out))
            np.testing.assert_allclose(out[0], G["bxdf_pdf"][i, j], rtol=1e-5, atol=1e-7, err_msg=f"pdf {d[:4]}")
            L.or_bxdf(p(d), 2, p(wo), p(wi), float(UC[i, j]), p(u), p(out))
            ref = G["bxdf_sample"][i, j]
            assert out[10] == ref[10], f"sample validity {d[:4]}"
            assert out[8] == ref[8], f"sample flags {d[:4]}"
            np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6, err_msg=f"sample {d[:4]}")
            worst = max(worst, float(np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1e-3))))
    print(f"bxdf worst relative deviation {worst:.2e}")


def test_coated_and_sheen_closures_bit_exact():
    """Layered (coated diffuse / conductor, clear and scattering slabs) and sheen closures: f, pdf
    and sample under both transport modes and all sample flags, bit for bit. The layered walk's
    random stream is hashed from the directions, so equal bits also pin the walk's draw order."""
    D = G["cl_desc"]
    for i in range(len(D)):
        fl = orc.closure(D[i], 3, G["cl_wo"][i, 0])
        assert np.array_equal(fl[:2], G["cl_flags"][i]), (i, fl[:2], G["cl_flags"][i])
        for j in range(G["cl_wo"].shape[1]):
            args = (G["cl_wo"][i, j], G["cl_wi"][i, j], float(G["cl_uc"][i, j]), G["cl_u"][i, j],
                    int(G["cl_mode"][i, j]), int(G["cl_rt"][i, j]))
            f = orc.closure(D[i], 0, *args)[:4]
            assert np.array_equal(f.view(np.uint32), G["cl_f"][i, j].view(np.uint32)), ("f", i, j, f, G["cl_f"][i, j])
            pdf = orc.closure(D[i], 1, *args)[0]
            assert np.float32(pdf).view(np.uint32) == G["cl_pdf"][i, j].view(np.uint32), ("pdf", i, j)
            s = orc.closure(D[i], 2, *args)
            assert np.array_equal(s.view(np.uint32), G["cl_sample"][i, j].view(np.uint32)), ("sample", i, j, s,
                                                                                            G["cl_sample"][i, j])
    # the fixture exercises both outcomes of sampling, proportional pdfs and nonzero layered values
    assert 0 < G["cl_sample"][..., 10].mean() < 1 and G["cl_sample"][..., 11].sum() > 10
    assert np.count_nonzero(G["cl_f"][:7].max(axis=2)) > 40


def test_sheen_albedo_table():
    for a_i, a in enumerate(G["sheen_alpha"]):
        vals = np.array([L.or_sheen_albedo(float(c), float(a)) for c in G["sheen_cos"]], np.float32)
        assert np.array_equal(vals, G["sheen_albedo"][a_i]), a


def test_fresnel_and_phase_function():
    for c, e, k, rd, rc in zip(G["fr_cos"], G["fr_eta"], G["fr_k"], G["fr_dielectric"], G["fr_complex"]):
        np.testing.assert_allclose(L.or_fr_dielectric(float(c), float(e)), rd, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(L.or_fr_complex(float(abs(c)), float(e), float(k)), rc, rtol=1e-5, atol=1e-8)
    for gi, g in enumerate(G["hg_g"]):
        vals = [L.or_hg(float(c), float(g)) for c in G["fr_cos"][:20]]
        np.testing.assert_allclose(vals, G["hg_val"][gi], rtol=1e-6)
        wi, pdf = np.zeros(3, np.float32), np.zeros(1, np.float32)
        L.or_sample_hg(p(G["hg_wo"]), float(g), p(f32(G["hg_u"][gi])), p(wi), p(pdf))
        np.testing.assert_allclose(wi, G["hg_sample"][gi, :3], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(pdf[0], G["hg_sample"][gi, 3], rtol=1e-5)


def test_filters():
    x = G["filter_x"]
    for t, r in ((0, 0.5), (1, 1.5), (2, 1.5)):
        out, nrm = np.zeros(len(x), np.float32), np.zeros(1, np.float32)
        L.or_filter(t, r, p(x), len(x), p(out), p(nrm))
        np.testing.assert_allclose(out, G["filter_vals"][t], rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(nrm[0], G["filter_norms"][t], rtol=1e-6)


@pytest.fixture(scope="module")
def c1_small(host):
    s = host.generate(1, 64, 64)
    return s, orc.OracleScene(s)


def test_c1_camera_and_world_sphere(c1_small):
    s, o = c1_small
    np.testing.assert_allclose(o.world_sphere(), G["c1_world_sphere"], rtol=1e-6)
    for (fx, fy), ref in zip(G["c1_cam_film"], G["c1_cam_rays"]):
        np.testing.assert_allclose(o.camera_ray(float(fx), float(fy)), ref, rtol=1e-6, atol=1e-7)


def test_c1_bvh_hits_bit_exact(c1_small):
    _, o = c1_small
    hits = o.trace(G["c1_rays"])
    ref = G["c1_hits"]
    assert np.array_equal(hits["prim"], ref["prim"])
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    m = ref["prim"] >= 0
    np.testing.assert_array_equal(hits["b1"][m], ref["b1"][m])
    np.testing.assert_array_equal(hits["b2"][m], ref["b2"][m])
    anyh = o.trace(G["c1_rays_any"], any_hit=True)
    assert np.array_equal(anyh["prim"] >= 0, G["c1_any"]["prim"] >= 0)
    print(f"closest hits {m.sum()}/{len(m)}, any-hit occluded {(anyh['prim'] >= 0).sum()}")


def test_c1_light_bvh_sampling_and_pmf(c1_small):
    s, o = c1_small
    P, N, U = G["c1_lp"], G["c1_ln"], G["c1_lu"]
    for i in range(len(P)):
        li, pmf = o.light_sample(P[i], N[i], float(U[i]))
        assert li == G["c1_light"][i], i
        np.testing.assert_allclose(pmf, G["c1_light_pmf"][i], rtol=1e-5, atol=1e-9)
        each = [o.light_pmf(P[i], N[i], k) for k in range(len(s.lights))]
        np.testing.assert_allclose(each, G["c1_pmf_each"][i], rtol=1e-5, atol=1e-9)
    pw = np.array([L.or_light_power(o.h, k) for k in range(len(s.lights))], np.float64)
    np.testing.assert_allclose(pw / pw.sum(), G["c1_power_pmf"], rtol=1e-5)


def test_c1_area_light_sample_li(c1_small):
    s, o = c1_small
    P, N = G["c1_lp"], G["c1_ln"]
    for i in range(len(P)):
        k = i % len(s.lights)
        lam, pdf = sample_visible(G["c1_lsl"][i])
        out = np.zeros(15, np.float32)
        L.or_light_sample_li(o.h, k, p(f32(P[i])), p(f32(N[i])), p(f32(G["c1_lsu"][i])), p(lam), p(pdf), p(out))
        ref = G["c1_lsi"][i]
        assert out[14] == ref[8], i
        if ref[8]:
            np.testing.assert_allclose(out[:8], ref[:8], rtol=2e-5, atol=1e-7, err_msg=f"sample {i}")
            got = L.or_light_pdf_li(o.h, k, p(f32(P[i])), p(f32(ref[4:7])))
            np.testing.assert_allclose(got, G["c1_lpdf"][i], rtol=2e-5)



def test_c5_thin_lens_camera_rays(host):
    """Configuration 5's f/2.8 thin-lens camera (src/cameras/camera.cpp:36-48) against the reference's
    rays for the same film positions and lens samples."""
    s = host.generate(5, 160, 90, 0.25)
    o = orc.OracleScene(s)
    for (fx, fy), (lx, ly), ref in zip(G["c5_cam_film"], G["c5_cam_lens"], G["c5_cam_rays"]):
        np.testing.assert_allclose(o.camera_ray(float(fx), float(fy), float(lx), float(ly)), ref, rtol=2e-6, atol=2e-6)
    # the lens matters: the centre sample and a rim sample of one film position leave different points
    a = o.camera_ray(80.0, 45.0, 0.5, 0.5)
    b = o.camera_ray(80.0, 45.0, 1.0, 0.5)
    assert np.linalg.norm(a[:3] - b[:3]) > 0.1 * s.camera.lensRadius


def test_c2_light_bvh_sampling_and_pmf(host):
    """The light BVH over configuration 2's 128 emitting triangles (src/lights/light_bvh.cpp:103-234):
    chosen light, its PMF and every light's PMF at 200 points, and the power PMFs."""
    s = host.generate(2, 160, 90, 0.1)
    o = orc.OracleScene(s)
    assert len(s.lights) == 128
    P, N, U = G["c2l_p"], G["c2l_n"], G["c2l_u"]
    for i in range(len(P)):
        li, pmf = o.light_sample(P[i], N[i], float(U[i]))
        assert li == G["c2l_light"][i], i
        np.testing.assert_allclose(pmf, G["c2l_light_pmf"][i], rtol=1e-5, atol=1e-9)
        each = [o.light_pmf(P[i], N[i], k) for k in range(len(s.lights))]
        np.testing.assert_allclose(each, G["c2l_pmf_each"][i], rtol=1e-5, atol=1e-9)
    pw = np.array([L.or_light_power(o.h, k) for k in range(len(s.lights))], np.float64)
    np.testing.assert_allclose(pw / pw.sum(), G["c2l_power_pmf"], rtol=1e-5)


def scene_file(host, name):
    s = host.load_scene(os.path.join(os.path.dirname(__file__), "scenes", name, "scene.prscene.json"))
    return s, orc.OracleScene(s)


@pytest.mark.parametrize("name,prefix,kind,count", [("spots", "spot", pa.LIGHT_SPOT, 3),
                                                     ("ies", "ies", pa.LIGHT_GONIO, 3),
                                                     ("sun", "sun", pa.LIGHT_SUN, 1)])
def test_spot_and_goniometric_lights_match_reference(host, name, prefix, kind, count):
    """Spot lights (reference src/lights/light.cpp:117-212: smoothstep falloff between the cone's
    cosines, cone bounds in the light BVH, PowerY) and goniometric lights (src/lights/sky.cpp:10-50 with
    the type C profile of src/lights/ies.cpp: the lookup, its symmetry folding, the tabulated integral)
    and the sun (src/lights/sky.cpp:82-127: a uniform cone over the disk, an infinite light in the
    light BVH, the attenuated solar spectrum) next to point, spot, environment and area lights:
    light-BVH choice and every light's PMF, power PMFs, SampleLi and PDF_Li against the
    reference's own code. The reference rebuilds each IES profile with its own
    LM-63 parser, so the power PMFs also pin the host's integral of the profile."""
    s, o = scene_file(host, name)
    kinds = s.lights["kind"]
    assert (kinds == kind).sum() == count
    P, N, U = G[f"{prefix}_p"], G[f"{prefix}_n"], G[f"{prefix}_u"]
    for i in range(len(P)):
        li, pmf = o.light_sample(P[i], N[i], float(U[i]))
        assert li == G[f"{prefix}_light"][i], i
        np.testing.assert_allclose(pmf, G[f"{prefix}_light_pmf"][i], rtol=1e-5, atol=1e-9)
        each = [o.light_pmf(P[i], N[i], k) for k in range(len(s.lights))]
        np.testing.assert_allclose(each, G[f"{prefix}_pmf_each"][i], rtol=1e-5, atol=1e-9)
    pw = np.array([L.or_light_power(o.h, k) for k in range(len(s.lights))], np.float64)
    np.testing.assert_allclose(pw / pw.sum(), G[f"{prefix}_power_pmf"], rtol=1e-5)
    lit = 0
    for i in range(len(P)):
        k = i % len(s.lights)
        lam, pdf = sample_visible(G[f"{prefix}_lsl"][i])
        out = np.zeros(15, np.float32)
        L.or_light_sample_li(o.h, k, p(f32(P[i])), p(f32(N[i])), p(f32(G[f"{prefix}_lsu"][i])), p(lam), p(pdf),
                             p(out))
        ref = G[f"{prefix}_lsi"][i]
        assert out[14] == ref[8], i
        if ref[8]:
            lit += kinds[k] == kind
            # 1e-4: inside a spot's falloff band the smoothstep turns the few-ulp differences of the
            # cone cosine (the reference normalises the stored unit axis once more) into up to ~150x
            # that relative change of the radiance
            np.testing.assert_allclose(out[:8], ref[:8], rtol=1e-4, atol=1e-7, err_msg=f"sample {i}")
            got = L.or_light_pdf_li(o.h, k, p(f32(P[i])), p(f32(ref[4:7])))
            np.testing.assert_allclose(got, G[f"{prefix}_lpdf"][i], rtol=2e-5, err_msg=f"pdf {i}")
    assert lit > 50  # the new light type is sampled inside and across its profile


def test_image_texture_lookup_bit_exact():
    """ImageData::Lookup with the texture's UVMapping (src/materials/texture.cpp:8-36): every wrap
    mode, with and without the v flip, two uv mappings, texel centres and edges and points far
    outside [0, 1]^2 -- the record's texels read back bit for bit as the reference reads its Image."""
    uv = G["tex_uv"]
    for case, rec, ref in zip(G["tex_cases"], G["tex_recs"], G["tex_ref"]):
        out = np.zeros_like(ref)
        L.or_tex_lookup(p(rec), p(uv), len(uv), p(out))
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), f"lookup {case}"


def test_shading_perturbation_bit_exact():
    """Material::PerturbShading (src/materials/material.cpp:16-58): normal maps (strength, flipped
    green), bump maps (mean or one channel, multiplier, scale) and both together, on random shading
    contexts -- the perturbed shading normal and tangent, and whether the frame changed."""
    knots, ctx = G["shade_knots"], G["shade_ctx"]
    for k, (sh, ref) in enumerate(zip(G["shade_recs"], G["shade_ref"])):
        out = np.zeros_like(ref)
        L.or_perturb_shading(p(knots), p(sh), p(ctx), len(ctx), p(out))
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), f"shading record {k}"
        assert ref[:, 6].all()
