import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for sub in ("viennacl-dev_amd", "oracle"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libvclhip")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.set_threads(1)  # the reference's single-thread summation order
    return pyoracle


@pytest.fixture(scope="session")
def vh():
    import vclhip
    return vclhip


@pytest.fixture(scope="session")
def dev(vh):
    """The HIP device; fails (does not skip) when none is visible."""
    n = vh.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    vh.call("vclhip_set_device", 0)
    return 0
