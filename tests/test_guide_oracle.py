"""The FP64 path-guiding model (oracle/guide_oracle.cpp) against what the mathematics fixes:
the analytic gradient against central finite differences, the vMF mixture's normalisation and
its sampler against its pdf. CPU only."""
import numpy as np

import oracle as orc

N_PARAMS = 4 * (64 * 64 + 64) + 48 * 64 + 48


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(N_PARAMS) * scale / 8).astype(np.float32)
    return P


def random_records(n, seed):
    rng = np.random.default_rng(seed)
    R = np.zeros((n, 16), np.float32)
    R[:, 0:3] = rng.random((n, 3))
    for a, b in ((3, 6), (6, 9), (9, 12)):
        v = rng.standard_normal((n, 3))
        R[:, a:b] = v / np.linalg.norm(v, axis=1, keepdims=True)
    R[:, 12] = np.abs(R[:, 11]) / np.pi  # a cosine-lobe pdf
    R[:, 13] = rng.random(n) * 2
    return R


def test_parameter_layout():
    assert orc.guide_param_count() == N_PARAMS
    # the meaningful parameters: 4 x 64 x 64 + 64 x 41 weights and 4 x 64 + 41 biases
    assert 4 * 64 * 64 + 64 * 41 + 4 * 64 + 41 == 19305


def test_gradient_matches_finite_differences():
    P = random_params(1, scale=2.0)
    R = random_records(12, 2)
    loss, g = orc.guide_loss_grad(P, R)
    assert np.isfinite(loss) and np.any(g != 0)
    rng = np.random.default_rng(3)
    # layer offsets: weights then biases, rows 64,64,64,64,48
    offs, o = [], 0
    for rows in (64, 64, 64, 64, 48):
        offs.append((o, rows))
        o += rows * 64 + rows
    idx = []
    for (o, rows) in offs:
        idx += list(o + rng.integers(0, rows * 64, 30))  # weights
        idx += list(o + rows * 64 + rng.integers(0, rows, 6))  # biases
    o4 = offs[4][0]
    idx += [o4 + 40 * 64 + i for i in range(0, 64, 7)] + [o4 + 48 * 64 + 40]  # selection row and bias
    checked = kinks = 0
    for i in idx:
        h = np.float32(2.0 ** -14)
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp, l0, lm = orc.guide_loss_grad(Pp, R)[0], loss, orc.guide_loss_grad(Pm, R)[0]
        fwd, bwd = (lp - l0) / (float(Pp[i]) - float(P[i])), (l0 - lm) / (float(P[i]) - float(Pm[i]))
        if abs(fwd - bwd) > 1e-2 * max(abs(fwd), abs(bwd)) + 1e-6:
            kinks += 1  # the step crosses a ReLU kink: the loss is not differentiable there
            continue
        fd = (lp - lm) / (float(Pp[i]) - float(Pm[i]))
        if abs(fd) < 1e-6 and abs(g[i]) < 1e-6:
            continue
        assert abs(fd - g[i]) <= 2e-3 * max(abs(fd), abs(g[i])) + 1e-6, (i, fd, g[i])
        checked += 1
    assert checked > 100 and kinks < 10, (checked, kinks)


def test_mixture_pdf_normalised_and_sampler_consistent():
    rng = np.random.default_rng(5)
    dirs = rng.standard_normal((20000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for trial in range(3):
        out = (rng.standard_normal(41) * 1.5).astype(np.float32)
        out[3:40:5] += 2.0 * trial  # sharper lobes in later trials
        p = np.array([orc.guide_mix_pdf(out, d)[0] for d in dirs])
        est = 4 * np.pi * p.mean()
        err = 4 * np.pi * p.std() / np.sqrt(len(p))
        assert abs(est - 1) < 4 * err + 0.01, (trial, est, err)
        # E[1 / pdf(w)] over w drawn from the mixture is the sphere's area
        inv = []
        for _ in range(4000):
            w = orc.guide_mix_sample(out, *rng.random(3).astype(np.float32))
            assert abs(np.linalg.norm(w) - 1) < 1e-5
            inv.append(1 / orc.guide_mix_pdf(out, w)[0])
        inv = np.array(inv)
        est = inv.mean()
        err = inv.std() / np.sqrt(len(inv))
        assert abs(est - 4 * np.pi) < 4 * err + 0.02 * 4 * np.pi, (trial, est, err)
    _, sel = orc.guide_mix_pdf(np.zeros(41, np.float32), np.array([0, 0, 1], np.float32))
    assert abs(sel - 0.5) < 1e-12
