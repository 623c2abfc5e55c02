"""GPU: qo_filter_records_noalloc (static scratch, no allocation) against
qo_filter_records (per-stream workspace) through the C ABI: two calls in a row,
the second on a smaller input, every output bit-identical (order included),
plus the size limit rejected with an error before anything runs."""
import ctypes as C

import numpy as np
import pytest

from query_on_opencl_fpga_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def _records(orc, n, seed):
    y = orc.gen_uniform_u32(n, seed, 0, 0, 1 << 31)
    return np.stack([np.arange(n, dtype=np.uint32), y], 1)


def _run(qo, fn, d_in, n, lo, hi, cap, offset_bytes=0):
    out = DeviceBuffer(max(cap, 1) * 8)
    cnt = DeviceBuffer(8)
    try:
        qo._lib.call("qo_fill_u32", out.ptr, 0xDEADBEEF, max(cap, 1) * 2, None)
        qo._lib.call(fn, C.c_void_p(d_in.ptr + offset_bytes), n, lo, hi, out.ptr, cnt.ptr, None)
        qo._lib.call("qo_stream_sync", None)
        m = int(cnt.to_numpy(np.uint64, 1)[0])
        return m, out.to_numpy(np.uint32, 2 * m).reshape(-1, 2)
    finally:
        out.free()
        cnt.free()


def test_noalloc_matches_filter_records_on_two_calls(qo, orc):
    lib = qo._lib.load()
    lib.qo_filter_records_noalloc.restype = C.c_int
    lib.qo_filter_records_noalloc.argtypes = [C.c_void_p, C.c_size_t, C.c_int32, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p]
    big, small = 5_000_011, 70_003
    rec = _records(orc, big, 31)
    d = DeviceBuffer.from_numpy(rec)
    try:
        lo, hi = 0, (1 << 30) - 1
        # Call 1: the whole input; call 2: a smaller, 8-byte-offset (unaligned) input
        # right after it, on the array the first call left behind.
        for n, off in ((big, 0), (small, 8)):
            m_ref, ref = _run(qo, "qo_filter_records", d, n, lo, hi, n, off)
            m_na, na = _run(qo, "qo_filter_records_noalloc", d, n, lo, hi, n, off)
            exp = orc.filter_records(rec[off // 8: off // 8 + n], lo, hi)
            assert m_na == m_ref == len(exp)
            assert np.array_equal(na, ref) and np.array_equal(na, exp)
    finally:
        d.free()


def test_noalloc_rejects_inputs_beyond_its_scratch(qo):
    lib = qo._lib.load()
    lib.qo_filter_records_noalloc.restype = C.c_int
    lib.qo_filter_records_noalloc.argtypes = [C.c_void_p, C.c_size_t, C.c_int32, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p]
    cnt = DeviceBuffer(8)
    try:
        # 2^31 + 1 records would need 2^17 + 1 tiles: rejected before any access.
        rc = lib.qo_filter_records_noalloc(C.c_void_p(16), (1 << 31) + 1, 0, 1, C.c_void_p(16), cnt.ptr, None)
        assert rc == 1 and "scratch" in qo._lib.last_error()
    finally:
        cnt.free()
