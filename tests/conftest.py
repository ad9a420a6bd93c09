import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (run with -m gpu on an MI355X)")


def _ensure_built():
    lib = os.path.join(ROOT, "cabanamd_amd", "libcabanamd_amd.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                       check=True)
    if not os.path.exists(lib):
        from cabanamd_amd import build
        build.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import cabanamd_amd as M
    if M.device_count() < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    return M
