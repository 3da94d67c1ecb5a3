"""Path-tracer parity of the MI355X wavefront kernels against the CPU oracle.

(ii) Per-sample spectral radiance: for the same (pixel, sample index, seed) both sides draw the
     same PCG32 stream (src/integrators/integrator.h:29-38), so every camera sample's 4-wavelength
     radiance must agree within 1e-4 relative. A sample whose path took a different branch (a
     Russian-roulette or Fresnel decision within float round-off of its threshold; libm and the
     GPU's math library differ in the last ulp) is counted as diverged; the diverged fraction is
     reported and bounded.
(iii) Converged images: the GPU and the oracle render the same scene with independent sample
     indices; per pixel and channel the means must agree within 3 sigma (sigma from each side's
     FP64 sum-of-squares accumulators), with the outlier fraction bounded binomially.
"""
import numpy as np
import pytest

import oracle as orc
import prender_amd as pa

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def radiance_agreement(gpu, cpu):
    """Fraction of samples whose 4 wavelengths agree within REL_TOL (absolute floor 1e-6)."""
    Lg, Lc = gpu["L"].astype(np.float64), cpu["L"].astype(np.float64)
    rel = np.abs(Lg - Lc) / np.maximum(np.abs(Lc), 1e-6)
    ok = np.all(rel <= REL_TOL, axis=1)
    return ok, rel


def check_samples(name, scene, params, plugin, min_agree):
    o = orc.OracleScene(scene)
    gpu = plugin.sample_radiance(scene, params)
    cpu = o.sample_radiance(params)
    # The camera sample itself (wavelengths, film position) is a pure function of the RNG stream.
    assert np.array_equal(gpu["pFilm"].view(np.uint32), cpu["pFilm"].view(np.uint32)), "film positions differ"
    np.testing.assert_allclose(gpu["lam"], cpu["lam"], rtol=2e-6, atol=0)
    ok, rel = radiance_agreement(gpu, cpu)
    nz = np.any(cpu["L"] != 0, axis=1)
    frac = ok.mean()
    print(f"{name}: samples={len(ok)} nonzero={nz.mean():.3f} agree={frac:.5f} diverged={int((~ok).sum())} "
          f"median rel err={np.median(rel[ok]):.2e}", flush=True)
    assert frac >= min_agree, f"{name}: only {frac:.4f} of samples agree within {REL_TOL}"
    return gpu, cpu


def test_sample_radiance_c1(plugin, host):
    s = host.generate(1, 96, 96)
    for si in (0, 7, 31):
        check_samples(f"c1 si={si}", s, pa.render_params(spp=1, first_sample=si, seed=1, max_depth=8, rr_start=3),
                      plugin, 0.995)


def test_sample_radiance_c2(plugin, host):
    s = host.generate(2, 160, 90)
    check_samples("c2", s, pa.render_params(spp=1, first_sample=3, seed=1, max_depth=16, rr_start=1), plugin, 0.99)


def test_sample_radiance_power_light_sampler(plugin, host):
    s = host.generate(2, 96, 54)
    p = pa.render_params(spp=1, first_sample=5, seed=9, max_depth=16, light_sampler=pa.LIGHT_SAMPLER_POWER)
    check_samples("c2-power", s, p, plugin, 0.99)


def splat_correlation(scene, n=4096, seed=0):
    """Variance inflation of an image mean over independent pixels when every sample is splatted
    into its filter footprint: E[(sum_k w_k)^2] / E[sum_k w_k^2] over sample positions."""
    cam = scene.camera
    r = 0.5 if cam.filterType == pa.FILTER_BOX else cam.filterRadius

    def w(x):
        x = np.abs(x)
        if cam.filterType == pa.FILTER_GAUSSIAN:
            sg = r / 3
            v = np.maximum(0, np.exp(-x * x / (2 * sg * sg)) - np.exp(-r * r / (2 * sg * sg)))
        elif cam.filterType == pa.FILTER_BH:
            t = 0.5 + x / (2 * r)
            v = 0.35875 - 0.48829 * np.cos(2 * np.pi * t) + 0.14128 * np.cos(4 * np.pi * t) - 0.01168 * np.cos(6 * np.pi * t)
        else:
            v = np.ones_like(x)
        return np.where(x > r, 0, v)

    u = np.random.default_rng(seed).random((n, 2))
    k = np.arange(-3, 4)
    sums = []
    for axis in range(2):
        wx = w(u[:, axis:axis + 1] - (k[None, :] + 0.5))
        sums.append((wx.sum(1), (wx * wx).sum(1)))
    return float(np.mean((sums[0][0] * sums[1][0]) ** 2) / np.mean(sums[0][1] * sums[1][1]))


def three_sigma(name, g_sum, g_sq, n_g, c_sum, c_sq, n_c, corr=1.0):
    mg, mc = g_sum / n_g, c_sum / n_c
    vg = np.maximum(g_sq / n_g - mg * mg, 0) / n_g
    vc = np.maximum(c_sq / n_c - mc * mc, 0) / n_c
    sigma = np.sqrt(vg + vc)
    diff = np.abs(mg - mc)
    valid = sigma > 0
    out = (diff > 3 * sigma + 1e-12) & valid
    frac = out.sum() / max(1, valid.sum())
    # Expected outlier fraction ~0.27%; allow a binomial 5-sigma margin and heavy-tail slack.
    n = valid.sum()
    bound = 0.0027 + 5 * np.sqrt(0.0027 * (1 - 0.0027) / max(1, n)) + 0.01
    # The image means, against the standard error of their difference; `corr` inflates the
    # per-pixel variance sum for the correlation the film's filter splat puts between neighbours.
    z = abs(mg.mean() - mc.mean()) / max(np.sqrt(corr * (vg + vc).sum()) / mg.size, 1e-30)
    rel_mean = abs(mg.mean() - mc.mean()) / mc.mean()
    print(f"{name}: pixels*channels={n} outside 3 sigma={frac:.4%} (bound {bound:.2%}) mean ratio diff={rel_mean:.3e} "
          f"(z = {z:.2f})", flush=True)
    assert frac <= bound, f"{name}: {frac:.4%} of pixel values outside 3 sigma"
    assert z < 5, f"{name}: image means differ by {z:.1f} standard errors"
    return frac


def test_image_3sigma_c1(plugin, host):
    s = host.generate(1, 48, 48)
    spp = 256
    pg = pa.render_params(spp=spp, first_sample=0, seed=1, max_depth=8, rr_start=3)
    g_sum, g_sq = plugin.render_ex(s, pg)
    pc = pa.render_params(spp=spp, first_sample=spp, seed=2, max_depth=8, rr_start=3)
    c_sum, c_sq, _ = orc.OracleScene(s).render(pc)
    three_sigma("c1 image", g_sum, g_sq, spp, c_sum, c_sq, spp, splat_correlation(s))


def test_film_matches_oracle_film_same_samples(plugin, host):
    """With identical sample indices the two films differ only where a sample diverged."""
    s = host.generate(1, 64, 64)
    p = pa.render_params(spp=4, first_sample=0, seed=1, max_depth=8, rr_start=3)
    g_sum, g_sq = plugin.render_ex(s, p)
    c_sum, c_sq, _ = orc.OracleScene(s).render(p)
    rel = np.abs(g_sum - c_sum) / np.maximum(np.abs(c_sum), 1e-3)
    close = np.mean(rel <= 1e-3)
    print(f"film same-samples: fraction within 1e-3 = {close:.4f}", flush=True)
    assert close >= 0.97
    np.testing.assert_allclose(g_sum.sum(), c_sum.sum(), rtol=5e-3)


def test_checkpoint_resume_equals_uninterrupted(plugin, host, tmp_path):
    """A render stopped at a checkpoint and resumed at its pass count equals the uninterrupted
    render: sample streams are indexed by sample number (reference gpu_path.cpp:462-474)."""
    s = host.generate(1, 48, 32)
    handle = host.open_plugin()
    p = pa.render_params(spp=1, first_sample=0, seed=1, max_depth=8, rr_start=3)
    full, full_sq, done, _ = host.render(handle, s, p, 12)
    assert done == 12
    part, part_sq, done1, _ = host.render(handle, s, p, 5)
    path = tmp_path / "c1.ckpt"
    host.checkpoint_write(path, s, 8, 1, part, part_sq, done1)
    film, sq, passes = host.checkpoint_read(path, s, 8, 1)
    p2 = pa.render_params(spp=1, first_sample=passes, seed=1, max_depth=8, rr_start=3)
    rest, rest_sq, done2, _ = host.render(handle, s, p2, 12 - passes)
    np.testing.assert_allclose(film + rest, full, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(sq + rest_sq, full_sq, rtol=1e-9, atol=1e-12)
    host.lib.ph_plugin_close(handle)


def test_unprofiled_bounce_loop_matches_profiled_rounds(plugin, host):
    """Unprofiled renders queue each bounce one ahead of the path-count read-back (or replay
    captured bounce graphs when those are compiled in); profiled renders synchronise after every
    bounce. Both must give the same film bit for bit, on every repetition, with other work and
    synchronisations in between (a plain scene, media with interfaces, coated materials)."""
    scenes = [host.generate(2, 96, 54, 0.25), host.generate(3, 64, 36, 0.125), host.generate(6, 64, 36, 0.25)]
    for k, scene in enumerate(scenes):
        prm = pa.render_params(spp=3, first_sample=5, seed=2, max_depth=24, rr_start=1)
        plugin.profile(1)
        eager, eager_sq = plugin.render_ex(scene, prm)
        plugin.profile_read(reset=True)
        plugin.profile(0)
        for rep in range(3):
            g, g_sq = plugin.render_ex(scene, prm)
            assert np.array_equal(g, eager) and np.array_equal(g_sq, eager_sq), f"scene {k} replay {rep}"
            plugin.profile(1)  # eager work between the replays
            plugin.render_ex(scene, pa.render_params(spp=1, first_sample=rep, seed=3, max_depth=4))
            plugin.profile(0)
