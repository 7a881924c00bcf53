import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "opencl-sph_amd")
for p in (PKG, os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def have_reference():
    return os.path.isdir(os.path.join(REFERENCE, "src"))


@pytest.fixture(scope="session")
def hip_device():
    """Initialise the HIP device once; every gpu test depends on it."""
    import sph
    sph.init_hip()
    return True
