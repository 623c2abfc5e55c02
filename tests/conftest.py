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


@pytest.fixture(autouse=True)
def _tune_table_restored(request, tmp_path_factory):
    """Every GPU test leaves libqo's launch-parameter table as it found it
    (entries and absent entries alike: saved to a file, then reset and
    reloaded), so parameters a test sets cannot leak into later tests."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    pkg = request.getfixturevalue("qo")
    snap = tmp_path_factory.mktemp("tune") / "table.tune"
    pkg._lib.call("qo_tune_save", str(snap).encode())
    yield
    pkg._lib.call("qo_tune_reset")
    pkg._lib.call("qo_tune_load", str(snap).encode())
