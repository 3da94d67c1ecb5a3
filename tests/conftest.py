import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ar-renderer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP plugin)")


@pytest.fixture(scope="session")
def host():
    import prender_amd as pa
    return pa.Host()


@pytest.fixture(scope="session")
def plugin():
    import prender_amd as pa
    p = pa.Plugin()
    devs = p.devices()
    if not devs:
        raise RuntimeError("GPU test requested but the plugin sees no HIP device")
    print("device:", devs[0].name.decode(), flush=True)
    return p


@pytest.fixture(scope="session")
def c1(host):
    return host.generate(1)
