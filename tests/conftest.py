"""Shared test setup: repo root on sys.path, the `gpu` marker, fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def device():
    """The HIP device handle.  GPU tests fail (not skip) without a device."""
    import gjxl_amd

    dev = gjxl_amd.Device(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def ref():
    """The reference's own CPU code (oracle/_ref), or skip where unbuilt."""
    from oracle import binding

    lib = binding.reference()
    if lib is None:
        pytest.skip("oracle/_ref is not built (reference sources absent)")
    return lib
