"""GPU: operators called from two host threads at once on the NULL stream (the
library's per-thread stream).  Each thread has its own workspace and counter
block, so the calls do not disturb each other; every result is checked
against the oracle."""
import threading

import numpy as np
import pytest

from query_on_opencl_fpga_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def test_two_threads_null_stream(qo, orc):
    n = 3_000_001
    inputs = []
    for t in range(2):
        y = orc.gen_uniform_u32(n, 40 + t, 0, 0, 1 << 31)
        x = orc.gen_uniform_u32(n, 40 + t, 1, 0, 1000)
        rec = np.stack([x, y], 1).astype(np.uint32)
        k, v = y.astype(np.int32), x.astype(np.int32)
        lo, hi = 0, (1 << 29) * (t + 1) - 1
        inputs.append((rec, k, v, lo, hi, orc.filter_records(rec, lo, hi), orc.filter_agg(k, v, lo, hi)))
    errors = []

    def worker(t):
        rec, k, v, lo, hi, exp_rec, exp_agg = inputs[t]
        d_rec = DeviceBuffer.from_numpy(rec)
        d_out = DeviceBuffer(rec.nbytes)
        d_cnt = DeviceBuffer(8)
        d_k, d_v, d_agg = DeviceBuffer.from_numpy(k), DeviceBuffer.from_numpy(v), DeviceBuffer(24)
        try:
            for _ in range(12):
                qo._lib.call("qo_filter_records", d_rec.ptr, n, lo, hi, d_out.ptr, d_cnt.ptr, None)
                qo._lib.call("qo_filter_agg", d_k.ptr, d_v.ptr, n, lo, hi, d_agg.ptr, None)
                qo._lib.call("qo_stream_sync", None)
                m = int(d_cnt.to_numpy(np.uint64, 1)[0])
                if m != len(exp_rec) or not np.array_equal(d_out.to_numpy(np.uint32, 2 * m).reshape(-1, 2), exp_rec):
                    errors.append(f"thread {t}: filter_records differs")
                    return
                if tuple(int(a) for a in d_agg.to_numpy(np.int64, 3)) != exp_agg:
                    errors.append(f"thread {t}: filter_agg differs")
                    return
        except Exception as e:  # surfaced below
            errors.append(f"thread {t}: {e!r}")
        finally:
            for b in (d_rec, d_out, d_cnt, d_k, d_v, d_agg):
                b.free()

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in ths), "a worker thread did not finish"
    assert not errors, errors
