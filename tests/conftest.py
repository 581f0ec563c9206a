import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")
    config.addinivalue_line("markers", "slow: long CPU case")


def _ensure_oracle():
    lib = ROOT / "oracle" / "liboracle_fdtd.so"
    if not lib.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def hip_ready():
    """The HIP library, loaded; fails (not skips) when it is missing on a GPU run."""
    from rfnetwork_amd import _lib

    L = _lib.lib()
    if not _lib.gpu_visible():
        pytest.fail("no HIP device visible for a gpu-marked test")
    return L
