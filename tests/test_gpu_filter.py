"""GPU parity: generators, fused filter+aggregate, stable filter compaction and
the exclusive scan against the oracle (bit-exact)."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KA = json.loads((Path(__file__).resolve().parent / "golden" / "known_answers.json").read_text())
INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1


def dev_gen(qo, name, n, *args):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    with DeviceBuffer(max(n, 1) * 4) as b:
        qo._lib.call(name, b.ptr, n, *args, None)
        return b.to_numpy(np.uint32, n)


def test_generators_match_oracle(qo, orc):
    n = 300_001
    assert np.array_equal(dev_gen(qo, "qo_gen_uniform_u32", n, 5, 2, 0, 1000), orc.gen_uniform_u32(n, 5, 2, 0, 1000))
    assert np.array_equal(dev_gen(qo, "qo_gen_uniform_u32", n, 5, 3, 7, 0), orc.gen_uniform_u32(n, 5, 3, 7, 0))
    for m in (1, 1000, 1 << 16, 70_001):
        assert np.array_equal(dev_gen(qo, "qo_gen_perm_u32", m, 77), orc.gen_perm(m, 77))
    b = dev_gen(qo, "qo_gen_bij_u32", 5000, 31, 9)
    assert b.tolist() == [orc.bij(i, 31, 9) for i in range(5000)]
    fk = dev_gen(qo, "qo_gen_fk_u32", 20_000, 4, 1, 5000, 0, 77)
    u = orc.gen_uniform_u32(20_000, 4, 1, 0, 5000)
    assert fk.tolist() == [orc.perm(int(x), 5000, 77) for x in u]
    G, z = 4096, 0.8
    t = np.empty(G, np.uint64)
    qo._lib.call("qo_zipf_table", t.ctypes.data, G, z)
    assert np.array_equal(t, orc.zipf_table(G, z))
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    with DeviceBuffer.from_numpy(t) as dt:
        zk = dev_gen(qo, "qo_gen_zipf_u32", 100_000, dt.ptr, G, 3, 4, 5)
    assert np.array_equal(zk, orc.gen_zipf_keys(100_000, t, 3, 4, 5))


def test_filter_agg_known_answer(qo):
    c = KA["filter_agg"]
    assert qo.operators.filter_agg(c["key"], c["val"], c["lo"], c["hi"]) == (c["count"], c["sum"], c["sum32"])


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1000, 4099, 1 << 20, (1 << 20) + 3])
def test_filter_agg_matches_oracle(qo, orc, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 1 << 31, n, dtype=np.int64).astype(np.int32)
    v = rng.integers(0, 1000, n, dtype=np.int32)
    lo, hi = 0, (1 << 30) - 1
    assert qo.operators.filter_agg(k, v, lo, hi) == orc.filter_agg(k, v, lo, hi)


def test_filter_agg_edge_predicates(qo, orc):
    rng = np.random.default_rng(5)
    n = 100_000
    k = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    v = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    for lo, hi in ((INT_MIN, INT_MAX), (1, 0), (0, 0), (INT_MIN, -1), (-5, 5)):
        assert qo.operators.filter_agg(k, v, lo, hi) == orc.filter_agg(k, v, lo, hi)


def test_filter_agg_unaligned_columns(qo, orc):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    rng = np.random.default_rng(6)
    n = 50_003
    k = rng.integers(INT_MIN, INT_MAX, n + 1, dtype=np.int64).astype(np.int32)
    v = rng.integers(0, 1000, n + 1, dtype=np.int32)
    with DeviceBuffer.from_numpy(k) as dk, DeviceBuffer.from_numpy(v) as dv, DeviceBuffer(24) as out:
        qo._lib.call("qo_filter_agg", dk.ptr + 4, dv.ptr + 4, n, -100, 1 << 29, out.ptr, None)
        got = tuple(int(x) for x in out.to_numpy(np.int64, 3))
    assert got == orc.filter_agg(k[1:], v[1:], -100, 1 << 29)


def test_filter_agg_records_matches_oracle(qo, orc):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    rng = np.random.default_rng(7)
    for n in (0, 1, 2, 3, 1 << 20, 12345):
        r = np.stack([rng.integers(0, 1000, n), rng.integers(0, 1 << 31, n)], 1).astype(np.uint32)
        with DeviceBuffer.from_numpy(r) as dr, DeviceBuffer(24) as out:
            qo._lib.call("qo_filter_agg_records", dr.ptr, n, 0, (1 << 30) - 1, out.ptr, None)
            got = tuple(int(x) for x in out.to_numpy(np.int64, 3))
        assert got == orc.filter_agg_records(r, 0, (1 << 30) - 1)


def test_filter_records_known_answers(qo):
    for c in KA["filter"]:
        got = qo.operators.filter_records(np.array(c["in"], np.uint32), c["lo"], c["hi"])
        assert got.tolist() == c["out"]


@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 100_000, (1 << 20) + 7])
@pytest.mark.parametrize("sel", [0.0, 0.06, 0.5, 1.0])
def test_filter_records_matches_oracle(qo, orc, n, sel):
    rng = np.random.default_rng(n)
    r = np.stack([np.arange(n), rng.integers(0, 100, n)], 1).astype(np.uint32)
    hi = int(sel * 100) - 1
    got = qo.operators.filter_records(r, 0, hi)
    assert np.array_equal(got, orc.filter_records(r, 0, hi))


def test_filter_records_unaligned(qo, orc):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    rng = np.random.default_rng(8)
    n = 30_001
    r = np.stack([np.arange(n + 1), rng.integers(-50, 50, n + 1)], 1).astype(np.int64).astype(np.uint32)
    with DeviceBuffer.from_numpy(r) as d, DeviceBuffer(n * 8) as o, DeviceBuffer(8) as c:
        qo._lib.call("qo_filter_records", d.ptr + 8, n, -10, 10, o.ptr, c.ptr, None)
        m = int(c.to_numpy(np.uint64, 1)[0])
        got = o.to_numpy(np.uint32, 2 * m).reshape(m, 2)
    assert np.array_equal(got, orc.filter_records(r[1:], -10, 10))


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 4096, 4097, 1 << 20, (1 << 22) + 5])
def test_scan_matches_oracle(qo, orc, n):
    rng = np.random.default_rng(n)
    x = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    assert np.array_equal(qo.operators.scanImpl(x), orc.scan_excl_i32(x))


def test_scan_known_answers_and_in_place(qo, orc):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    for c in KA["scan"]:
        assert qo.operators.scanImpl(c["in"]).tolist() == c["out"]
    x = np.ones(1_000_003, np.int32)
    with DeviceBuffer.from_numpy(x) as d:
        qo._lib.call("qo_scan_exclusive_i32", d.ptr, d.ptr, x.shape[0], None)
        assert np.array_equal(d.to_numpy(np.int32, x.shape[0]), np.arange(x.shape[0], dtype=np.int32))


def test_scan_unaligned_pointers(qo, orc):
    """Columns 4 bytes off a 16-byte boundary take the one-element-per-load path
    (four times as many tiles, and look-back status words, as the aligned path)."""
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    n = 3_000_001
    rng = np.random.default_rng(11)
    x = rng.integers(-1000, 1000, n + 1, dtype=np.int64).astype(np.int32)
    with DeviceBuffer.from_numpy(x) as d, DeviceBuffer((n + 1) * 4) as o:
        qo._lib.call("qo_scan_exclusive_i32", d.ptr + 4, o.ptr + 4, n, None)
        got = o.to_numpy(np.int32, n + 1)[1:]
    assert np.array_equal(got, orc.scan_excl_i32(x[1:]))
