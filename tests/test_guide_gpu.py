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


def test_sample_radiance_room(plugin, c4):
    check_samples("c4", c4, pa.render_params(spp=1, first_sample=3, seed=1, max_depth=32), plugin, 0.99)


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
    three_sigma("c4 guided vs plain oracle", g_sum, g_sq, spp, c_sum, c_sq, spp, splat_correlation(c4))
