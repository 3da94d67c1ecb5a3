"""Host integration on the CPU: the render session (resume, checkpoints, progress, cancel, time limit,
errors; reference src/integrators/gpu_path.cpp:440-513 and src/integrators/integrator.h:49-95) and the
CLI's JSON events (src/cli/main.cpp:299-305, 453-468, 576-587).

The session drives a stand-in plugin (tests/fake_plugin/fake_plugin.cpp) whose film after samples
[0, n) is known in closed form: pixel i holds (i + 1) n (n + 1) / 2 in every channel of the sum and
(i + 1)^2 n (n + 1) (2n + 1) / 6 in the sum of squares. The GPU counterpart of the resume check runs the
real plugin in test_session_gpu.py.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import prender_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def fake_plugin_path(tmp_path_factory):
    out = tmp_path_factory.mktemp("fake") / "libfake_prgpu.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(HERE, "fake_plugin", "fake_plugin.cpp"), "-o", str(out)], check=True)
    return str(out)


@pytest.fixture(scope="module")
def fake(host, fake_plugin_path):
    h = host.open_plugin(fake_plugin_path)
    yield h
    host.lib.ph_plugin_close(h)


@pytest.fixture(scope="module")
def tiny(host):
    return host.generate(1, 8, 4)


@pytest.fixture
def env(monkeypatch):
    monkeypatch.delenv("FAKE_PRGPU_SLEEP_MS", raising=False)
    monkeypatch.delenv("FAKE_PRGPU_FAIL_AT", raising=False)
    return monkeypatch


def film_after(scene, n):
    i = np.arange(1, scene.width * scene.height + 1, dtype=np.float64).reshape(scene.height, scene.width, 1)
    s = np.repeat(i * (n * (n + 1) / 2), 3, axis=2)
    q = np.repeat(i * i * (n * (n + 1) * (2 * n + 1) / 6), 3, axis=2)
    return s, q


def dumps(obj):
    return json.dumps(obj, sort_keys=True, separators=(",", ":"), ensure_ascii=False)


# ---- JSON events ---------------------------------------------------------------------------------

def test_devices_event_lists_cpu_then_plugin_devices(host, fake_plugin_path):
    line = host.list_devices_json(fake_plugin_path, cpu_threads=12)
    expected = {"event": "devices", "gpu_status": "ok", "devices": [
        {"id": "cpu", "name": "CPU", "backend": "native", "memory_mb": 0.0, "compute_units": 12, "gpu": False},
        {"id": "gpu:0", "name": 'Fake "GPU" 0', "backend": "fake", "memory_mb": 1536.0, "compute_units": 7,
         "gpu": True}]}
    assert line == dumps(expected)
    assert line.startswith('{"devices":[{"backend":"native","compute_units":12,"gpu":false,"id":"cpu",')


def test_devices_event_reports_loader_failure(host, tmp_path):
    ev = json.loads(host.list_devices_json(str(tmp_path / "missing.so"), cpu_threads=3))
    assert [d["id"] for d in ev["devices"]] == ["cpu"]
    assert ev["gpu_status"].startswith("cannot load GPU plugin")
    default = json.loads(host.list_devices_json(None))
    assert default["devices"][0]["compute_units"] == os.cpu_count()


# nlohmann::json dump() of a double: shortest round-trip digits; fixed notation while the decimal
# point position n (value = 0.digits x 10^n) lies in (-4, 15], with ".0" on integral values;
# d.ddde+XX (two exponent digits at least) outside it; 0 as 0.0; NaN and infinities as null.
NUMBER_CASES = [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-2.5, "-2.5"), (16.0, "16.0"), (0.1, "0.1"),
    (1 / 3, "0.3333333333333333"), (123.456, "123.456"), (1e-4, "0.0001"), (1.5e-4, "0.00015"),
    (1e-5, "1e-05"), (2.5e-7, "2.5e-07"), (1e14, "100000000000000.0"), (123456789012345.0, "123456789012345.0"),
    (1e15, "1e+15"), (1.5e15, "1.5e+15"), (1e16, "1e+16"), (1.5e300, "1.5e+300"), (5e-324, "5e-324"),
    (1.7976931348623157e308, "1.7976931348623157e+308"), (float("nan"), "null"), (float("inf"), "null"),
    (float("-inf"), "null"),
]


@pytest.mark.parametrize("value,text", NUMBER_CASES)
def test_progress_event_number_format(host, value, text):
    line = host.progress_json("render", value, 0.5, 2.0)
    assert line == '{"elapsed":0.5,"event":"progress","fraction":%s,"samples_per_pixel":2.0,"stage":"render"}' % text


def test_event_numbers_agree_with_python_json_outside_the_e15_decade(host):
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.random(300), 10.0 ** rng.uniform(-30, 30, 300), -(10.0 ** rng.uniform(-8, 14, 100))])
    for v in vals:
        if 1e15 <= abs(v) < 1e16:
            continue
        got = json.loads(host.progress_json("render", float(v), 0.0, 0.0))
        assert got["fraction"] == v
        assert host.progress_json("render", float(v), 0.0, 0.0) == dumps(
            {"event": "progress", "stage": "render", "fraction": float(v), "elapsed": 0.0, "samples_per_pixel": 0.0})


def test_progress_accept_rate_stage_and_error_events(host):
    assert host.progress_json("mlt", 0.25, 1.0, 3.0, 0.5) == dumps(
        {"event": "progress", "stage": "mlt", "fraction": 0.25, "elapsed": 1.0, "samples_per_pixel": 3.0,
         "accept_rate": 0.5})
    assert host.stage_json("render") == '{"event":"stage","name":"render"}'
    msg = 'bad "scene"\\ at\tline\n2 \x01 — ü'
    line = host.error_json(msg)
    assert line == dumps({"event": "error", "message": msg})
    assert "\\u0001" in line and "ü" in line
    # a short buffer is truncated and NUL-terminated; the full length is returned
    buf = C.create_string_buffer(8)
    n = host.lib.ph_stage_json(b"render", buf, 8)
    assert n == len('{"event":"stage","name":"render"}') and buf.value == b'{"event'


# ---- render session ------------------------------------------------------------------------------

def session(host, fake, scene, spp, **kw):
    params = pa.render_params(seed=kw.pop("seed", 7), max_depth=kw.pop("max_depth", 5))
    return host.render_session(fake, scene, params, spp, **kw)


def test_session_full_job_progress_and_batches(host, fake, tiny, env):
    calls = []
    s, q, done, scale = session(host, fake, tiny, 37, progress_interval=0.0,
                                on_progress=lambda *a: calls.append(a) and False)
    es, eq = film_after(tiny, 37)
    assert done == 37 and scale == 1 / 37
    assert np.array_equal(s, es) and np.array_equal(q, eq)
    # batches double from 1 while a batch takes < 0.15 s: 1, 2, 4, 8, 16, then the remaining 6
    assert [c[3] for c in calls] == [1, 3, 7, 15, 31, 37]
    assert all(c[0] == "render" for c in calls)
    assert [c[1] for c in calls] == [n / 37 for n in (1, 3, 7, 15, 31, 37)]
    assert all(b[2] >= a[2] >= 0 for a, b in zip(calls, calls[1:]))


def test_session_progress_is_throttled_but_reports_the_last_batch(host, fake, tiny, env):
    calls = []
    session(host, fake, tiny, 37, progress_interval=3600.0, on_progress=lambda *a: calls.append(a) and False)
    assert [c[3] for c in calls] == [37] and calls[0][1] == 1.0


def test_session_cancel_checkpoint_and_resume(host, fake, tiny, env, tmp_path):
    ck = tmp_path / "job.ckpt"
    s, q, done, scale = session(host, fake, tiny, 37, checkpoint=ck, progress_interval=0.0,
                                on_progress=lambda st, f, e, n: n >= 7)
    assert done == 7 and scale == 1 / 7
    es, eq = film_after(tiny, 7)
    assert np.array_equal(s, es) and np.array_equal(q, eq)
    # the checkpoint written on cancel holds the film and the sample count
    cs, cq, passes = host.checkpoint_read(ck, tiny, 5, 7)
    assert passes == 7 and np.array_equal(cs, es) and np.array_equal(cq, eq)
    # resuming continues from sample 7 and ends with the film of an uninterrupted job
    calls = []
    s2, q2, done2, _ = session(host, fake, tiny, 37, resume=ck, checkpoint=ck, progress_interval=0.0,
                               on_progress=lambda *a: calls.append(a) and False)
    es, eq = film_after(tiny, 37)
    assert done2 == 37 and np.array_equal(s2, es) and np.array_equal(q2, eq)
    assert [c[3] for c in calls] == [8, 10, 14, 22, 37]
    assert host.checkpoint_read(ck, tiny, 5, 7)[2] == 37


def test_session_writes_periodic_checkpoints_after_each_due_batch(host, fake, tiny, env, tmp_path):
    ck = tmp_path / "periodic.ckpt"
    seen = []

    def on_progress(stage, fraction, elapsed, n):
        # the checkpoint of the previous batch is on disk when the next batch reports
        seen.append((n, host.checkpoint_read(ck, tiny, 5, 7)[2] if ck.exists() else None))
        return False

    session(host, fake, tiny, 15, checkpoint=ck, checkpoint_interval=0.0, progress_interval=0.0,
            on_progress=on_progress)
    assert seen == [(1, None), (3, 1), (7, 3), (15, 7)]
    assert host.checkpoint_read(ck, tiny, 5, 7)[2] == 15


def test_session_cancel_flag_set_before_the_start(host, fake, tiny, env, tmp_path):
    ck = tmp_path / "never.ckpt"
    flag = C.c_int32(1)
    s, q, done, scale = session(host, fake, tiny, 37, checkpoint=ck, cancel=flag)
    assert done == 0 and scale == 0.0 and not s.any()
    assert host.checkpoint_read(ck, tiny, 5, 7)[2] == 0


def test_session_time_limit(host, fake, tiny, env):
    env.setenv("FAKE_PRGPU_SLEEP_MS", "20")
    calls = []
    s, q, done, scale = session(host, fake, tiny, 0, time_limit=0.25, progress_interval=0.0,
                                on_progress=lambda *a: calls.append(a) and False)
    assert 3 <= done < 1 << 20 and scale == 1 / done
    assert calls[-1][1] == 1.0 and calls[-1][2] >= 0.25
    assert all(c[1] == min(1.0, c[2] / 0.25) for c in calls)
    es, _ = film_after(tiny, done)
    assert np.array_equal(s, es)


def test_session_errors(host, fake, tiny, env, tmp_path):
    env.setenv("FAKE_PRGPU_FAIL_AT", "5")
    with pytest.raises(RuntimeError, match=r"^GPU render failed: injected failure at sample 5$"):
        session(host, fake, tiny, 37)
    env.delenv("FAKE_PRGPU_FAIL_AT")
    ck = tmp_path / "a.ckpt"
    session(host, fake, tiny, 3, checkpoint=ck)
    with pytest.raises(RuntimeError, match="was made for a different scene, integrator, device, resolution, "
                                           "seed or depth"):
        session(host, fake, tiny, 8, resume=ck, seed=8)
    with pytest.raises(RuntimeError, match="was made for a different"):
        session(host, fake, tiny, 8, resume=ck, max_depth=6)
    with pytest.raises(RuntimeError, match="cannot open checkpoint"):
        session(host, fake, tiny, 8, resume=tmp_path / "missing.ckpt")
    # a failed checkpoint write is a warning, not a failed job
    s, q, done, _ = session(host, fake, tiny, 3, checkpoint=tmp_path / "no" / "such" / "dir.ckpt")
    assert done == 3
    # a plugin without the extension is refused
    with pytest.raises(RuntimeError, match="lacks the extension interface"):
        host.render_session(None, tiny, pa.render_params(), 1)

