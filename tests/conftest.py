import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libqo kernels on the device)")


@pytest.fixture(scope="session")
def qo():
    """The product package; GPU tests fail loudly if libqo.so is missing."""
    import query_on_opencl_fpga_amd as pkg

    pkg._lib.call("qo_init", 0)
    return pkg


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle
