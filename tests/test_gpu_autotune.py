"""GPU: the autotuner searches on the device and keeps a candidate in the table."""
import ctypes as C
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op,cls,choices", [(0, 20, {(2,), (4,), (8,)}), (8, 12, {(1,), (2,), (4,), (8,)}),
                                            (5, 20, {(16,), (24,), (32,), (48,), (60,)})])
def test_autotune_sets_a_candidate(qo, op, cls, choices):
    lib = qo._lib
    b, it, pa = C.c_int(), C.c_int(), C.c_int()
    best, dflt = C.c_double(), C.c_double()
    try:
        lib.call("qo_autotune", op, cls, 2, C.byref(b), C.byref(it), C.byref(pa), C.byref(best), C.byref(dflt))
        assert (it.value,) in choices and best.value > 0 and dflt.value > 0
        gb, gi, gp = C.c_int(), C.c_int(), C.c_int()
        lib.call("qo_tune_get", op, cls, C.byref(gb), C.byref(gi), C.byref(gp))
        assert (gb.value, gi.value, gp.value) == (b.value, it.value, pa.value)
    finally:
        # Back to the table libqo loaded at init (the committed tune file).
        lib.call("qo_tune_reset")
        arch = qo.device.device_info()["arch"].split(":")[0]
        f = Path(lib.LIB_PATH).parent / "tune" / f"{arch}.tune"
        if f.exists():
            lib.call("qo_tune_load", str(f).encode())
