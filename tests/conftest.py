import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "build", "liblpbf_oracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle()


@pytest.fixture(scope="session")
def ora():
    from oracle import pyoracle
    return pyoracle


_GPU_ERROR_MARKERS = ("AcceleratorError", "hipError", "HIP error", "memory access fault",
                      "illegal memory access", "No HIP GPUs")


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    """Stop the session at the first GPU runtime error: after a device fault no
    further GPU work is started in this process."""
    outcome = yield
    rep = outcome.get_result()
    if rep.failed and call.excinfo is not None:
        text = f"{call.excinfo.typename}: {call.excinfo.value}"
        if any(m in text for m in _GPU_ERROR_MARKERS):
            item.session.shouldstop = f"GPU runtime error in {item.nodeid}: {text[:200]}"
