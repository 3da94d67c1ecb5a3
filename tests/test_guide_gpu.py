"""Learned path guiding on the MI355X plugin.

The BF16 matrix-core network (inference and the fused training kernel) is checked against the
FP64 model of oracle/guide_oracle.cpp on the same BF16-representable parameters: outputs and the
loss gradient agree to the precision BF16 activations allow (relative L2 error, stated below).
A guided render, which samples the learned one-sample MIS proposal and trains the model between
batches, must stay unbiased: its converged image matches the oracle's plain path tracer within
3 sigma on the indoor configuration.
"""
import numpy as np
import pytest

import oracle as orc
import prender_amd as pa
from test_guide_oracle import random_records
from test_pt_gpu import check_samples, splat_correlation, three_sigma

pytestmark = pytest.mark.gpu

FWD_TOL = 3e-2   # relative L2 error of the 41 outputs: BF16 inputs and activations, FP32 accumulation
GRAD_TOL = 8e-2  # relative L2 error of the gradient: BF16 activations and output gradients


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def model_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    n = orc.guide_param_count()
    P = (rng.standard_normal(n) * scale / 8).astype(np.float32)
    return bf16_round(P)


def test_guide_forward_matches_fp64_model(plugin):
    P = model_params(11)
    plugin.guide_params(P)
    np.testing.assert_array_equal(plugin.guide_params(), P)
    R = random_records(3000, 12)[:, :9]
    R[:, 0:3] = bf16_round(R[:, 0:3])
    gpu = plugin.guide_eval(R)
    cpu = orc.guide_forward(P, R)
    err = np.linalg.norm(gpu - cpu) / np.linalg.norm(cpu)
    print(f"guide forward: relative L2 error {err:.3e}", flush=True)
    assert err < FWD_TOL


def test_guide_gradient_matches_fp64_model(plugin):
    P = model_params(21)
    plugin.guide_params(P)
    R = random_records(4096, 22)
    g_gpu, st = plugin.guide_grad(R)
    loss, g_cpu = orc.guide_loss_grad(P, R)
    err = np.linalg.norm(g_gpu - g_cpu) / np.linalg.norm(g_cpu)
    cos = float(g_gpu @ g_cpu / (np.linalg.norm(g_gpu) * np.linalg.norm(g_cpu)))
    print(f"guide gradient: relative L2 error {err:.3e}, cosine {cos:.5f}; loss gpu {st[1] / st[0]:.5f} "
          f"cpu {loss:.5f}", flush=True)
    np.testing.assert_allclose(st[0], R[:, 13].sum(), rtol=1e-4)
    np.testing.assert_allclose(st[1] / st[0], loss, rtol=2e-2)
    assert err < GRAD_TOL and cos > 0.995


def test_guide_training_reduces_loss(plugin):
    """Adam on records whose directions concentrate around a position-dependent axis."""
    rng = np.random.default_rng(31)
    n = 1 << 14
    R = random_records(n, 32)
    axis = np.stack([np.cos(4 * R[:, 0]), np.sin(4 * R[:, 0]), 0.3 * np.ones(n)], 1)
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    d = axis + 0.25 * rng.standard_normal((n, 3))
    R[:, 9:12] = d / np.linalg.norm(d, axis=1, keepdims=True)
    R[:, 12] = 1 / (4 * np.pi)
    R[:, 13] = 1.0
    plugin.guide_params(bf16_round(model_params(33, 0.5)))
    first = plugin.guide_train(R, 1)
    last = plugin.guide_train(R, 300, lr=3e-3)
    l0, l1 = first[1] / first[0], last[1] / last[0]
    loss_cpu, _ = orc.guide_loss_grad(plugin.guide_params(), R)
    print(f"guide training: loss {l0:.4f} -> {l1:.4f} (FP64 model on the trained parameters {loss_cpu:.4f})",
          flush=True)
    assert l1 < l0 - 0.5
    assert abs(loss_cpu - l1) < 0.1 * abs(l1) + 0.05


@pytest.fixture(scope="module")
def c4(host):
    return host.generate(4, 96, 54, 0.25)


def look_at_camera(scene, eye, at, fov_y, width, height):
    """The scene with a pinhole camera at `eye` looking at `at` (camera x right, y up, z forward)."""
    eye, at = np.asarray(eye, np.float64), np.asarray(at, np.float64)
    f = at - eye
    f /= np.linalg.norm(f)
    r = np.cross(np.array([0.0, 1.0, 0.0]), f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    cam = pa.CameraEx.from_buffer_copy(bytes(scene.camera))
    m = np.stack([r, u, f, eye], axis=1).astype(np.float32)  # 3 x 4, columns: right, up, forward, origin
    for i, v in enumerate(m.reshape(-1)):
        cam.cameraToWorld[i] = float(v)
    cam.tanHalfFov = float(np.tan(np.radians(fov_y) / 2))
    cam.width, cam.height, cam.lensRadius = width, height, 0.0
    return pa.Scene(scene.prims, scene.materials, scene.spectra, scene.knots, scene.lights, scene.media, scene.grid, cam,
                    scene.desc.cameraMedium, scene.defaults)


def test_sample_radiance_room(plugin, c4):
    check_samples("c4", c4, pa.render_params(spp=1, first_sample=3, seed=1, max_depth=32), plugin, 0.99)


def test_sample_radiance_room_sunlit_furniture(plugin, c4):
    """A view along the sunlight of the furniture the sun lights through the window (the camera sits
    2 units towards the sun from a lit box top), so that most samples carry radiance: direct sun, and
    light bounced off the lit faces."""
    sun = np.array(c4.lights[c4.lights["kind"] == pa.LIGHT_DISTANT][0]["direction"], np.float64)
    at = np.array([2.54, 0.58, 2.41])
    s = look_at_camera(c4, at + 2.0 * sun, at, 20.0, 72, 48)
    _, cpu = check_samples("c4 sunlit furniture", s, pa.render_params(spp=1, first_sample=5, seed=1, max_depth=32),
                           plugin, 0.99)
    assert np.mean(np.any(cpu["L"] != 0, axis=1)) > 0.3


def test_guided_render_is_unbiased(plugin, c4):
    """Eight calls of 32 spp: the model is reset on the first and trained after every call."""
    spp, calls = 256, 8
    g_sum = g_sq = 0
    for k in range(calls):
        flags = pa.RENDER_GUIDE | pa.RENDER_GUIDE_TRAIN | (pa.RENDER_GUIDE_RESET if k == 0 else 0)
        p = pa.render_params(spp=spp // calls, first_sample=k * spp // calls, seed=1, max_depth=32, flags=flags)
        s, q = plugin.render_ex(c4, p)
        g_sum, g_sq = g_sum + s, g_sq + q
    c_sum, c_sq, _ = orc.OracleScene(c4).render(pa.render_params(spp=spp, first_sample=spp, seed=2, max_depth=32))
    corr = splat_correlation(c4)
    three_sigma("c4 guided vs plain oracle", g_sum, g_sq, spp, c_sum, c_sq, spp, corr)
    # the plain GPU render at the same budget, for scale: this scene's heavy tails alone give z of this size
    p_sum, p_sq = plugin.render_ex(c4, pa.render_params(spp=spp, first_sample=4 * spp, seed=3, max_depth=32))
    three_sigma("c4 plain vs plain oracle", p_sum, p_sq, spp, c_sum, c_sq, spp, corr)


def _image_mean_z(a_sum, a_sq, b_sum, b_sq, spp, corr):
    ma, mb = a_sum / spp, b_sum / spp
    va = np.maximum(a_sq / spp - ma * ma, 0) / spp
    vb = np.maximum(b_sq / spp - mb * mb, 0) / spp
    se = np.sqrt(corr * (va + vb).sum()) / ma.size
    return (ma.mean() - mb.mean()) / mb.mean(), (ma.mean() - mb.mean()) / se


def test_guided_mean_matches_plain_at_high_spp(plugin, c4):
    """Once trained, the guided estimator (a fixed model within every call) has the plain estimator's
    mean: the two GPU image means agree within 1% at 16 384 spp per pixel and within 0.5% at 65 536,
    and the difference measured in standard errors does not grow with the sample count (a bias of
    a few percent would be tens of standard errors at these counts)."""
    s = look_at_camera(c4, (-3.0, 2.5, -3.5), (2.0, 0.5, 2.5), 70.0, 48, 32)
    corr = splat_correlation(s)
    for k in range(8):
        flags = pa.RENDER_GUIDE | pa.RENDER_GUIDE_TRAIN | (pa.RENDER_GUIDE_RESET if k == 0 else 0)
        plugin.render_ex(s, pa.render_params(spp=64, first_sample=k * 64, seed=1, max_depth=32, flags=flags),
                         variance=False)
    zs = []
    for spp, tol in ((16384, 0.01), (65536, 0.005)):
        g = plugin.render_ex(s, pa.render_params(spp=spp, first_sample=1 << 20, seed=5, max_depth=32,
                                                 flags=pa.RENDER_GUIDE))
        q = plugin.render_ex(s, pa.render_params(spp=spp, first_sample=1 << 21, seed=6, max_depth=32))
        rel, z = _image_mean_z(g[0], g[1], q[0], q[1], spp, corr)
        print(f"c4 guided vs plain GPU at {spp} spp: image mean difference {rel:+.4%} (z = {z:+.2f})", flush=True)
        assert abs(rel) < tol
        zs.append(abs(z))
    assert zs[1] < 5


def _mixture_directions(out, rng):
    """One direction per row from the decoded 8-lobe vMF mixture of the network's 41 outputs (the
    decode of guide.h: lobe k's mean = normalised outputs 5k..5k+2, kappa = min(0.01 + softplus(o[5k+3]),
    1000), weights = softmax of o[5k+4])."""
    n = len(out)
    logit = out[:, 4:40:5]
    w = np.exp(logit - logit.max(1, keepdims=True))
    w /= w.sum(1, keepdims=True)
    mu = np.stack([out[:, 5 * k:5 * k + 3] for k in range(8)], 1)
    ln = np.linalg.norm(mu, axis=2, keepdims=True)
    mu = np.where(ln > 1e-6, mu / np.maximum(ln, 1e-30), np.array([0.0, 0.0, 1.0]))
    kap = np.minimum(0.01 + np.logaddexp(0, out[:, 3:40:5]), 1000.0)
    k = np.minimum((rng.random(n)[:, None] >= np.cumsum(w, 1)).sum(1), 7)
    m, kp = mu[np.arange(n), k], kap[np.arange(n), k]
    u1, u2 = rng.random(n), rng.random(n)
    ct = np.clip(1 + np.log1p(u1 * np.expm1(-2 * kp)) / kp, -1, 1)  # vMF cos theta about the mean
    return _around(m, ct, 2 * np.pi * u2)


def _around(axis, ct, phi):
    s = np.where(axis[:, 2:3] >= 0, 1.0, -1.0)[:, 0]
    a = -1 / (s + axis[:, 2])
    b = axis[:, 0] * axis[:, 1] * a
    t = np.stack([1 + s * axis[:, 0] ** 2 * a, s * b, -s * axis[:, 0]], 1)
    u = np.stack([b, s + axis[:, 1] ** 2 * a, -axis[:, 1]], 1)
    st = np.sqrt(np.maximum(0, 1 - ct * ct))
    return (st * np.cos(phi))[:, None] * t + (st * np.sin(phi))[:, None] * u + ct[:, None] * axis


def _sunlit(plugin, scene, sun, origins, dirs):
    """Whether the ray from each origin along dirs reaches a point the sun lights directly (the hit
    face turned to the sun and nothing in between), with the plugin's own ray queries."""
    rays = np.zeros(len(dirs), pa.RAY_DTYPE)
    rays["o"] = (origins + 1e-3 * dirs).astype(np.float32)
    rays["d"] = dirs.astype(np.float32)
    rays["tMax"] = np.inf
    h = plugin.trace(scene, rays)
    hit = h["prim"] >= 0
    q = scene.prims[h["prim"][hit]]
    nh = np.cross(q["p1"] - q["p0"], q["p2"] - q["p0"]).astype(np.float64)
    nh /= np.linalg.norm(nh, axis=1, keepdims=True)
    nh *= -np.sign(np.sum(nh * dirs[hit], 1))[:, None]
    ph = origins[hit] + dirs[hit] * h["t"][hit][:, None]
    facing = nh @ sun > 0
    s = np.zeros(int(facing.sum()), pa.RAY_DTYPE)
    s["o"] = (ph[facing] + 1e-3 * nh[facing]).astype(np.float32)
    s["d"] = sun.astype(np.float32)
    s["tMax"] = np.inf
    out = np.zeros(len(dirs), bool)
    out[np.flatnonzero(hit)[np.flatnonzero(facing)[plugin.trace(scene, s, any_hit=True)["prim"] < 0]]] = True
    return out


def test_trained_guide_sends_samples_toward_the_sunlight(plugin, host):
    """What the trained network gains over the sampling it is mixed with. The room is lit only by sun
    through a window, so indirect light comes from the few sunlit patches. After a short training
    render (8 calls of 16 spp at 160x90, the model trained on the GPU between calls), at the vertices
    of the last batch's training records, directions drawn from the learned mixture reach a sunlit
    point far more often than cosine-weighted directions about the same normals (the Lambertian
    BSDF's proposal). Measured: 14.6% against 1.0%; the bar is 6% and five times cosine."""
    import raygen
    scene = host.generate(4, 160, 90, 0.25)
    keys = ("guide_batch", "guide_min_steps", "guide_lr", "guide_bounces")
    saved = {k: plugin.tune_get(k) for k in keys}
    try:
        for k, v in zip(keys, (65536, 32, 3e-3, 1)):
            plugin.tune_set(k, v)
        plugin.film_reset(160, 90)
        for k in range(8):
            f = pa.RENDER_GUIDE | pa.RENDER_GUIDE_TRAIN | (pa.RENDER_GUIDE_RESET if k == 0 else 0)
            plugin.render_resident(scene, pa.render_params(spp=16, first_sample=16 * k, seed=1, max_depth=32,
                                                           rr_start=1, flags=f))
        R, produced = plugin.guide_records(1 << 20)
        rng = np.random.default_rng(5)
        sub = R[rng.choice(len(R), min(len(R), 20000), replace=False)]
        out = plugin.guide_eval(sub[:, :9]).astype(np.float64)
    finally:
        for k, v in saved.items():
            plugin.tune_set(k, v)
    assert len(R) > 5000, f"only {len(R)} training records"
    sun = np.array(scene.lights[scene.lights["kind"] == pa.LIGHT_DISTANT][0]["direction"], np.float64)
    sun /= np.linalg.norm(sun)
    lo, hi = raygen.scene_bounds(scene)
    pos = lo + sub[:, 0:3].astype(np.float64) * (hi - lo)
    nrm = sub[:, 3:6].astype(np.float64)
    nrm *= np.sign(np.sum(nrm * sub[:, 6:9], 1))[:, None]  # the side the path arrived from
    guided = _mixture_directions(out, rng)
    cosine = _around(nrm, np.sqrt(1 - rng.random(len(sub))), 2 * np.pi * rng.random(len(sub)))
    fg = _sunlit(plugin, scene, sun, pos, guided).mean()
    fc = _sunlit(plugin, scene, sun, pos, cosine).mean()
    print(f"trained guide: {len(R)} records ({produced} produced); directions reaching sunlit points: "
          f"mixture {fg:.4f}, cosine {fc:.4f}", flush=True)
    assert fg >= 0.06 and fg >= 5 * fc
