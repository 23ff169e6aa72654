import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """The oracle library is test infrastructure; build it if it is missing."""
    so = os.path.join(ROOT, "oracle", "libedoracle.so")
    if not os.path.exists(so):
        import fcntl
        with open(so + ".lock", "w") as lk:       # parallel test workers build it once
            fcntl.flock(lk, fcntl.LOCK_EX)
            if not os.path.exists(so):
                subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    yield


@pytest.fixture(scope="session")
def gpu_ready():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected without a visible GPU")
    import ed_code_amd
    ed_code_amd.init(0)
    return True
