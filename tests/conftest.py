import os
import subprocess
import sys

import pytest

try:  # share one HIP runtime with torch (see turboquant/_runtime.py)
    import torch  # noqa: F401
except ImportError:
    pass

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PKG_ROOT = os.path.join(REPO, "turboquant-space_amd")
for p in (PKG_ROOT, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950)")
    # The oracle is test infrastructure; build it here if build() has not.
    lib = os.path.join(REPO, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle")])


def gpu_available():
    try:
        from turboquant import capi
        return capi.lib().tq_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    from turboquant import capi
    n = capi.lib().tq_device_count()
    assert n > 0, "no GPU visible to libtq_hip.so"
    return capi
