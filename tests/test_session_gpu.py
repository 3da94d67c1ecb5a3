"""Render session and devices event with the HIP plugin (reference src/integrators/gpu_path.cpp:78-101,
440-513; src/cli/main.cpp:453-468). The CPU suite covers the session's bookkeeping with a stand-in
plugin; here the real renderer checks that a job cancelled at a checkpoint and resumed from it ends
with the film of the uninterrupted job."""
import json

import numpy as np
import pytest

import prender_amd as pa

pytestmark = pytest.mark.gpu


def test_devices_event_lists_the_hip_device(plugin, host):
    ev = json.loads(host.list_devices_json(pa.PLUGIN_PATH, cpu_threads=16))
    print("devices event:", json.dumps(ev), flush=True)
    assert ev["event"] == "devices" and ev["gpu_status"] == "ok"
    assert ev["devices"][0] == {"id": "cpu", "name": "CPU", "backend": "native", "memory_mb": 0.0,
                                "compute_units": 16, "gpu": False}
    infos = plugin.devices()
    assert [d["id"] for d in ev["devices"][1:]] == [f"gpu:{i}" for i in range(len(infos))]
    g = ev["devices"][1]
    assert g["gpu"] is True and g["name"] == infos[0].name.decode()
    assert g["memory_mb"] == infos[0].globalMemory / 1048576.0 and g["compute_units"] == infos[0].computeUnits


def test_session_cancel_and_resume_equals_uninterrupted(plugin, host, tmp_path):
    s = host.generate(1, 48, 32)
    handle = host.open_plugin()
    prm = pa.render_params(seed=3, max_depth=8, rr_start=3)
    full, full_sq, done, scale = host.render_session(handle, s, prm, 12)
    assert done == 12 and scale == 1 / 12
    ck = tmp_path / "session.ckpt"
    calls = []

    def stop_at_3(stage, fraction, elapsed, spp):
        calls.append(spp)
        return spp >= 3

    part, _, done1, _ = host.render_session(handle, s, prm, 12, checkpoint=ck, progress_interval=0.0,
                                            on_progress=stop_at_3)
    # batches double from 1 spp while they stay fast; the job stops at the first report >= 3
    assert 3 <= done1 < 12 and calls[-1] == done1
    rest, rest_sq, done2, scale2 = host.render_session(handle, s, prm, 12, resume=ck, checkpoint=ck)
    assert done2 == 12 and scale2 == 1 / 12
    print(f"resumed film vs uninterrupted: max rel diff "
          f"{np.max(np.abs(rest - full) / np.maximum(np.abs(full), 1e-12)):.2e}", flush=True)
    np.testing.assert_allclose(rest, full, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(rest_sq, full_sq, rtol=1e-9, atol=1e-12)
    assert np.any(part != full)
    host.lib.ph_plugin_close(handle)
