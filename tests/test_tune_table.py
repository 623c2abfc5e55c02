"""CPU: the launch-parameter table -- a tune file changes the parameters of the
classes it names, a class it does not name takes the nearest named class of
the same operator, and reset restores the built-in defaults (no GPU calls)."""
import ctypes as C
import pathlib

import pytest

LIB = pathlib.Path(__file__).resolve().parents[1] / "query_on_opencl_fpga_amd" / "libqo.so"
TUNE = pathlib.Path(__file__).resolve().parents[1] / "query_on_opencl_fpga_amd" / "tune"


def get(lib, op, cls):
    b, i, p = C.c_int(), C.c_int(), C.c_int()
    assert lib.qo_tune_get(op, cls, C.byref(b), C.byref(i), C.byref(p)) == 0
    return b.value, i.value, p.value


@pytest.mark.skipif(not LIB.exists(), reason="libqo.so is not built")
def test_tune_file_sets_classes_and_nearest_fallback(tmp_path):
    lib = C.CDLL(str(LIB))
    lib.qo_tune_load.argtypes = [C.c_char_p]
    lib.qo_tune_save.argtypes = [C.c_char_p]
    assert lib.qo_tune_reset() == 0
    default = get(lib, 0, 24)
    f = tmp_path / "t.tune"
    f.write_text("0 20 256 2 8\n0 28 256 8 1\n8 12 256 1 2048\n")
    assert lib.qo_tune_load(str(f).encode()) == 0
    assert get(lib, 0, 20) == (256, 2, 8)
    assert get(lib, 0, 28) == (256, 8, 1)
    assert get(lib, 0, 23) == (256, 2, 8)   # nearest: 20 (3 away) before 28 (5 away)
    assert get(lib, 0, 25) == (256, 8, 1)   # nearest: 28
    assert get(lib, 0, 24) == (256, 2, 8)   # tie: the smaller class
    assert get(lib, 8, 22) == (256, 1, 2048)
    assert get(lib, 3, 24) == get(lib, 3, 10)  # an operator the file does not name keeps its default
    out = tmp_path / "saved.tune"
    assert lib.qo_tune_save(str(out).encode()) == 0
    assert sorted(out.read_text().split("\n")) == sorted(f.read_text().split("\n"))
    assert lib.qo_tune_reset() == 0
    assert get(lib, 0, 24) == default


def test_committed_tune_file_is_well_formed():
    files = sorted(TUNE.glob("*.tune"))
    for f in files:
        for line in f.read_text().split("\n"):
            if line.strip():
                op, cls, b, i, p = map(int, line.split())
                assert 0 <= op < 10 and 0 <= cls < 64 and b > 0 and i > 0
