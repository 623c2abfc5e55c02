"""Reference-named entry points, exercised the way the reference's drivers do
(host/host/testFilter.cpp:197, testSort.cpp:287, testGroupBy.cpp:155,
testAggAfterGB.cpp:127, testScan.cpp:342) and checked with the reference's
validator conditions (host/host/Validate.cpp) plus the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def gen_records(orc, n, maxmax, seed):
    """Records with .x = row id and .y uniform in [0, maxmax) (the shape of
    generateRand, host/host/generator.cpp:63-77, with a portable generator)."""
    y = orc.gen_uniform_u32(n, seed, 0, 0, maxmax)
    return np.stack([np.arange(n, dtype=np.uint32), y], 1)


@pytest.mark.parametrize("n", [1, 1000, 1 << 16, 1 << 20])
def test_testFilterImpl(qo, orc, n):
    r = gen_records(orc, n, 100, 1)
    small = int(orc.gen_uniform_u32(1, 2, 0, 0, 100)[0])
    large = min(small + 6, 100)
    out = qo.operators.CL_RangeSelection(r, small, large)
    # validateFilter (Validate.cpp:317-370): count, then order-preserved rows
    sel = (r[:, 1].astype(np.int32) >= small) & (r[:, 1].astype(np.int32) <= large)
    assert out.shape[0] == int(sel.sum())
    assert np.array_equal(out, r[sel])
    assert np.array_equal(out, orc.filter_records(r, small, large))


@pytest.mark.parametrize("n", [2048, 1 << 16, 1 << 20, 5000])
def test_testSortImpl(qo, orc, n):
    r = gen_records(orc, n, 200, 3)
    out = qo.operators.CL_RadixSort(r)
    assert np.all(np.diff(out[:, 1].astype(np.int64)) >= 0)  # validateSort
    assert np.array_equal(out, orc.sort_records(r))          # plus permutation + stable tie rule


@pytest.mark.parametrize("n", [1, 1 << 12, 1 << 20])
def test_testGroupByImpl(qo, orc, n):
    r = gen_records(orc, n, 64, 4)
    srt, sp = qo.operators.CL_GroupBy(r)
    assert np.array_equal(srt, orc.sort_records(r))
    assert np.array_equal(sp, orc.group_startpos(srt))  # validateGroupBy: numGroups and startPos


@pytest.mark.parametrize("op", ["max", "min", "sum", "avg"])
def test_testAggAfterGroupByImpl(qo, orc, op):
    n = 1 << 18
    r = gen_records(orc, n, 50, 5)
    srt, sp = qo.operators.CL_GroupBy(r)
    fn = getattr(qo.operators, f"CL_agg_{op}_afterGroupBy")
    code = {"sum": 0, "max": 1, "min": 2, "avg": 3}[op]
    got = fn(srt, sp)
    assert np.array_equal(got, orc.parallel_aggregate(srt, sp, code))
    if op == "max":
        # validateAggAfterGroupBy_easy (Validate.cpp:72-164) indexes results by key value
        check = np.zeros(sp.shape[0], np.int64)
        np.maximum.at(check, srt[:, 1].astype(np.int64), srt[:, 0].astype(np.int32).astype(np.int64))
        assert np.array_equal(got, check)


def test_agg_with_gather(qo, orc):
    """aggAfterGroupByImpl with Ragg: S[i] = Ragg[Rin[i].x] is reduced (gatherBeforeAgg)."""
    n = 50_000
    r = gen_records(orc, n, 40, 6)
    srt, sp = qo.operators.CL_GroupBy(r)
    ragg = np.stack([orc.gen_uniform_u32(n, 7, 0, 0, 0), orc.gen_uniform_u32(n, 7, 1, 0, 0)], 1)
    gathered = ragg[srt[:, 0]]
    for code, fn in ((0, "sum"), (1, "max"), (2, "min"), (3, "avg")):
        got = getattr(qo.operators, f"CL_agg_{fn}_afterGroupBy")(srt, sp, ragg)
        assert np.array_equal(got, orc.parallel_aggregate(gathered, sp, code)), fn


def test_testScanImpl(qo, orc):
    n = (1 << 20) + 11
    x = orc.gen_uniform_u32(n, 8, 0, 0, n).astype(np.int32)
    got = qo.operators.scanImpl(x)
    assert np.array_equal(got, orc.scan_excl_i32(x))  # validateScan


def test_compat_errors_do_not_exit(qo):
    with pytest.raises(qo._lib.QoError):
        qo.operators.groupby_agg(np.arange(10, dtype=np.int32), np.arange(10, dtype=np.int32), 2)
    qo._lib.load().qo_debug_inject_fault(1)
    with pytest.raises(qo._lib.QoError) as ei:
        qo.operators.filter_agg(np.arange(10, dtype=np.int32), np.arange(10, dtype=np.int32), 0, 5)
    assert ei.value.status == 6
    assert qo.operators.filter_agg(np.arange(10, dtype=np.int32), np.arange(10, dtype=np.int32), 0, 5)[0] == 6


def _mem(qo, arr):
    lib = qo._lib.load()
    a = np.ascontiguousarray(arr)
    m = lib.qo_mem_create(max(a.nbytes, 4))
    if a.nbytes:
        qo._lib.call("qo_mem_write", m, a.ctypes.data, a.nbytes)
    return m


def _read(qo, m, dtype, count):
    out = np.empty(count, dtype)
    if out.nbytes:
        qo._lib.call("qo_mem_read", out.ctypes.data, m, out.nbytes)
    return out


@pytest.mark.parametrize("op", ["max", "min", "sum", "avg"])
def test_device_buffer_agg_afterGroupBy(qo, orc, op):
    """agg_{max,min,sum,avg}_afterGroupBy (AggAfterGroupBy.cpp:9-24) over cl_mem-style buffers."""
    lib = qo._lib.load()
    n = 100_003
    r = gen_records(orc, n, 300, 11)
    srt, sp = qo.operators.CL_GroupBy(r)
    g = sp.shape[0]
    d_r, d_sp = _mem(qo, srt), _mem(qo, sp.astype(np.int32))
    d_out = lib.qo_mem_create(g * 4)
    try:
        qo._lib.load().qo_clear_error()
        getattr(lib, f"agg_{op}_afterGroupBy")(d_r, n, d_sp, g, None, d_out, 256, None, None, None, None, None, 0)
        assert qo._lib.last_error() == ""
        got = _read(qo, d_out, np.int32, g)
        code = {"sum": 0, "max": 1, "min": 2, "avg": 3}[op]
        assert np.array_equal(got, orc.parallel_aggregate(srt, sp, code))
    finally:
        for m in (d_r, d_sp, d_out):
            lib.qo_mem_free(m)


def test_sort_only_variants(qo, orc):
    """CL_RadixSortOnly (Sort.cpp:11) sorts in place; CL_BitonicSortOnly (Sort.cpp:20)
    copies the input to d_Rout and sorts d_Rin in place; radixSortImpl (testSort.cpp:274)."""
    lib = qo._lib.load()
    n = 70_001
    r = gen_records(orc, n, 1 << 20, 12)
    exp = orc.sort_records(r)
    a = _mem(qo, r)
    lib.CL_RadixSortOnly(a, n, 256, 128, 0)
    assert np.array_equal(_read(qo, a, np.uint32, 2 * n).reshape(-1, 2), exp)
    b, out = _mem(qo, r), lib.qo_mem_create(r.nbytes)
    lib.CL_BitonicSortOnly(b, n, out, 256, 128, 0)
    assert np.array_equal(_read(qo, b, np.uint32, 2 * n).reshape(-1, 2), exp)
    assert np.array_equal(_read(qo, out, np.uint32, 2 * n).reshape(-1, 2), r)
    c = _mem(qo, r)
    lib.radixSortImpl(c, n, 32, 256, 128, None, None, None, None, None, 0)
    assert np.array_equal(_read(qo, c, np.uint32, 2 * n).reshape(-1, 2), exp)
    for m in (a, b, out, c):
        lib.qo_mem_free(m)


@pytest.mark.parametrize("n,offset", [(1, 0), (1000, 7), ((1 << 20) + 3, 1 << 16)])
def test_scan_offset_and_scanpara(qo, orc, n, offset):
    """scanImpl_offset (the offset overload of testScan.h:16) scans [offset, offset + n);
    initScan fills the reference's pass / gLength / blockSize (testScan.cpp:26-75)."""
    import ctypes as C

    lib = qo._lib.load()
    x = orc.gen_uniform_u32(n + offset, 13, 0, 0, 1 << 20).astype(np.int32)
    d_x, d_y = _mem(qo, x), lib.qo_mem_create(4 * n)
    sp = qo._lib.ScanPara()
    lib.initScan(n, C.byref(sp))
    try:
        assert sp.gLength == n and sp.blockSize == 256 and not sp.outputBuffer and not sp.tempBuffer
        p, v = 0, float(n)
        while v >= 256.0:
            v /= 256.0
            p += 1
        if n > 1 and p > 0 and 256 ** p == n:
            p -= 1
        assert sp.pass_ == p
        lib.scanImpl_offset(d_x, offset, n, d_y, None, None, None, None, None, C.byref(sp), 0)
        assert np.array_equal(_read(qo, d_y, np.int32, n), orc.scan_excl_i32(x[offset:offset + n]))
    finally:
        lib.closeScan(C.byref(sp))
        lib.qo_mem_free(d_x)
        lib.qo_mem_free(d_y)
