"""GPU: split, scatter, gather, projection, reduce, point selection and the
Record -> column step against the oracle, the brute-force fixtures and the
reference validators' conditions (host/host/Validate.cpp: validateSplit
:510-527, validateProjection :39-71, validateReduce :372-468), driven the way
the reference's drivers call them (testSplit.cpp:211-256, testScatter.cpp:
132-200)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from query_on_opencl_fpga_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

BF = json.loads((Path(__file__).resolve().parent / "golden" / "brute_force.json").read_text())


def rec(a):
    return np.array(a, dtype=np.uint32).reshape(-1, 2)


def gen_records(orc, n, seed, maxmax=0):
    """.x = row id, .y uniform (generateRand, host/host/generator.cpp:63-77)."""
    return np.stack([np.arange(n, dtype=np.uint32), orc.gen_uniform_u32(n, seed, 0, 0, maxmax)], 1)


# ---- split ------------------------------------------------------------------------------
@pytest.mark.parametrize("i", range(len(BF["split"])))
def test_split_fixtures(qo, i):
    c = BF["split"][i]
    out, bound = qo.operators.splitImpl(rec(c["in"]), c["parts"])
    assert out.tolist() == c["out"] and bound.tolist() == c["bound"]


@pytest.mark.parametrize("n,parts", [(1, 1), (4097, 2), (100_000, 32), (1 << 20, 255), (1 << 20, 256),
                                     ((1 << 20) + 5, 257), (300_001, 1000), (1 << 20, 65536)])
def test_split_matches_oracle(qo, orc, n, parts):
    r = gen_records(orc, n, 21 + parts)
    out, bound = qo.operators.splitImpl(r, parts)
    exp, exp_b = orc.split_records(r, parts)
    assert np.array_equal(out, exp)
    assert np.array_equal(bound, exp_b)
    pid = (out[:, 1] % parts).astype(np.int64)
    assert np.all(np.diff(pid) >= 0)  # validateSplit


def test_split_reference_driver_shape(qo, orc):
    """testSplitImpl(16M, 32, 128, 128) (host/host/main.cpp:564): generateRand
    with TEST_MAX, 32 partitions; validateSplit's condition plus the oracle."""
    n, parts = 16 * 1024 * 1024, 32
    r = gen_records(orc, n, 77, 1 << 30)
    out, bound = qo.operators.splitImpl(r, parts)
    pid = (out[:, 1] % parts).astype(np.int64)
    assert np.all(np.diff(pid) >= 0)
    exp, exp_b = orc.split_records(r, parts)
    assert np.array_equal(out, exp) and np.array_equal(bound, exp_b)


def test_split_native_without_bounds_and_errors(qo, orc):
    n, parts = 70_001, 300
    r = gen_records(orc, n, 5)
    exp, _ = orc.split_records(r, parts)
    with DeviceBuffer.from_numpy(r) as d_in, DeviceBuffer(n * 8) as d_out:
        qo._lib.call("qo_split_records", d_in.ptr, n, parts, d_out.ptr, None, None)
        assert np.array_equal(d_out.to_numpy(np.uint32, 2 * n).reshape(-1, 2), exp)
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_split_records", d_in.ptr, n, 65537, d_out.ptr, None, None)
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_split_records", d_in.ptr, n, 0, d_out.ptr, None, None)
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_split_records", d_in.ptr, n, 4, d_in.ptr, None, None)


# ---- gather / scatter ---------------------------------------------------------------------
@pytest.mark.parametrize("i", range(len(BF["gather"])))
def test_gather_fixtures(qo, i):
    c = BF["gather"][i]
    assert qo.operators.gatherImpl(rec(c["R"]), c["loc"], rec(c["S0"])).tolist() == c["S"]


@pytest.mark.parametrize("i", range(len(BF["scatter"])))
def test_scatter_fixtures_unique_targets(qo, i):
    """Fixture positions hit by exactly one in-range location must match; a
    position hit several times holds one of those records; others keep S0."""
    c = BF["scatter"][i]
    r, s0 = rec(c["R"]), rec(c["S0"])
    got = qo.operators.scatterImpl(r, c["loc"], s0)
    _check_scatter(r, np.array(c["loc"], np.int64), s0, got)


def _check_scatter(r, loc, s0, got):
    n = r.shape[0]
    ok = (loc >= 0) & (loc < n)
    hits = np.bincount(loc[ok], minlength=n) if n else np.zeros(0, np.int64)
    exp = s0.copy()
    uniq = ok & (hits[np.clip(loc, 0, max(n - 1, 0))] == 1) if n else ok
    exp[loc[uniq]] = r[uniq]
    single_or_none = hits <= 1
    assert np.array_equal(got[single_or_none], exp[single_or_none])
    dup = np.nonzero(hits > 1)[0]
    if dup.size:
        # Candidates of position p: the in-range sources with loc == p.
        src = np.nonzero(ok)[0]
        src = src[np.argsort(loc[src], kind="stable")]
        lo = np.searchsorted(loc[src], dup, "left")
        hi = np.searchsorted(loc[src], dup, "right")
        r64 = np.ascontiguousarray(r).view(np.uint64).ravel()
        g64 = np.ascontiguousarray(got).view(np.uint64).ravel()[dup]
        matched = np.zeros(dup.size, bool)
        for k in range(int((hi - lo).max())):
            live = lo + k < hi
            matched |= live & (r64[src[np.minimum(lo + k, src.size - 1)]] == g64)
        assert matched.all()


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 3, 1 << 24])
def test_gather_scatter_permutation(qo, orc, n):
    r = np.stack([orc.gen_uniform_u32(n, 3, 0, 0, 0), orc.gen_uniform_u32(n, 3, 1, 0, 0)], 1)
    p = orc.gen_perm(n, 17).astype(np.int32)
    g = qo.operators.gatherImpl(r, p)
    assert np.array_equal(g, r[p])
    assert np.array_equal(g, orc.gather_records(r, p))
    s = qo.operators.scatterImpl(r, p)
    assert np.array_equal(s, orc.scatter_records(r, p))
    assert np.array_equal(qo.operators.scatterImpl(g, p), r)  # scatter inverts gather


def test_scatter_gather_reference_driver_shape(qo, orc):
    """testScatterImpl / testGatherImpl (testScatter.cpp:132-200): loc uniform in
    [0, rLen) with repeats (generateRandInt), plus locations out of range."""
    n = 3_000_017
    r = gen_records(orc, n, 8, 1 << 30)
    loc = orc.gen_uniform_u32(n, 9, 0, 0, n).astype(np.int32)
    loc[::1000] = -5
    loc[7::1001] = n + 3
    s0 = np.stack([orc.gen_uniform_u32(n, 10, 0, 0, 0), orc.gen_uniform_u32(n, 10, 1, 0, 0)], 1)
    assert np.array_equal(qo.operators.gatherImpl(r, loc, s0), orc.gather_records(r, loc, s0))
    _check_scatter(r, loc.astype(np.int64), s0, qo.operators.scatterImpl(r, loc, s0))


def test_scatter_for_part_places_a_split(qo, orc):
    """scatterImpl_forPart with the split's own locations (stable counting-sort
    positions) yields the split."""
    n, parts = 200_003, 40
    r = gen_records(orc, n, 12)
    pid = r[:, 1] % parts
    order = np.argsort(pid, kind="stable")
    loc = np.empty(n, np.int32)
    loc[order] = np.arange(n, dtype=np.int32)
    lib = qo._lib.load()
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer.from_numpy(loc) as d_l, DeviceBuffer(n * 8) as d_s:
        m = [lib.qo_mem_wrap(b.ptr, b.nbytes) for b in (d_r, d_l, d_s)]
        lib.qo_clear_error()
        lib.scatterImpl_forPart(m[0], n, parts, m[1], m[2], 256, 128, None, None, None, None, None, 0)
        assert qo._lib.last_error() == ""
        for x in m:
            lib.qo_mem_free(x)
        assert np.array_equal(d_s.to_numpy(np.uint32, 2 * n).reshape(-1, 2), orc.split_records(r, parts)[0])


# ---- projection -----------------------------------------------------------------------------
@pytest.mark.parametrize("i", range(len(BF["project"])))
def test_projection_fixtures(qo, i):
    c = BF["project"][i]
    assert qo.operators.projection(rec(c["R"]), rec(c["proj"])).tolist() == c["out"]


def test_projection_matches_oracle(qo, orc):
    n, m = 1 << 22, 3_000_001
    r = gen_records(orc, n, 30)
    proj = np.stack([orc.gen_uniform_u32(m, 31, 0, 0, n + 100), orc.gen_uniform_u32(m, 31, 1, 0, 0)], 1)
    got = qo.operators.projection(r, proj)
    assert np.array_equal(got, orc.project_records(r, proj))
    inr = proj[:, 0] < n
    assert np.array_equal(got[inr, 1], r[proj[inr, 0], 1])  # validateProjection's rule


def test_projection_in_place_and_aliasing_errors(qo, orc):
    """out may overwrite proj (each row reads only its own proj row); gather into
    R itself and a projection writing R race, and are rejected."""
    n, m = 50_000, 70_001
    r = gen_records(orc, n, 32)
    proj = np.stack([orc.gen_uniform_u32(m, 33, 0, 0, n + 100), orc.gen_uniform_u32(m, 33, 1, 0, 0)], 1)
    loc = orc.gen_uniform_u32(n, 34, 0, 0, n).astype(np.int32)
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer.from_numpy(proj) as d_p, \
            DeviceBuffer.from_numpy(loc) as d_l:
        qo._lib.call("qo_project_records", d_r.ptr, n, d_p.ptr, m, d_p.ptr, None)
        assert np.array_equal(d_p.to_numpy(np.uint32, 2 * m).reshape(-1, 2), orc.project_records(r, proj))
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_project_records", d_r.ptr, n, d_p.ptr, m, d_r.ptr, None)
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_gather_records", d_r.ptr, n, d_l.ptr, n, d_r.ptr, None)
        assert np.array_equal(d_r.to_numpy(np.uint32, 2 * n).reshape(-1, 2), r)  # nothing ran


# ---- reduce ---------------------------------------------------------------------------------
@pytest.mark.parametrize("i", range(len(BF["reduce"])))
def test_reduce_fixtures(qo, i):
    c = BF["reduce"][i]
    assert [qo.operators.reduce(rec(c["in"]), op) for op in range(4)] == c["out"]


@pytest.mark.parametrize("n", [1, 2, 3, 4097, (1 << 20) + 1, 1 << 26])
def test_reduce_matches_oracle(qo, orc, n):
    r = gen_records(orc, n, 40 + n % 7)
    for op in range(4):
        assert qo.operators.reduce(r, op) == orc.reduce_records(r, op), op


def test_reduce_unaligned_and_repeated(qo, orc):
    """An 8-byte-aligned (not 16) input takes the scalar path; the same workspace
    serves calls in a row (its ticket is cleared by every call)."""
    n = 500_001
    r = gen_records(orc, n + 1, 41)
    with DeviceBuffer.from_numpy(r) as d, DeviceBuffer(4) as out:
        for op in (0, 1, 2, 3, 0):
            qo._lib.call("qo_reduce_records", C.c_void_p(d.ptr + 8), n, op, out.ptr, None)
            assert int(out.to_numpy(np.uint32, 1)[0]) == orc.reduce_records(r[1:], op)
        with pytest.raises(qo._lib.QoError):
            qo._lib.call("qo_reduce_records", d.ptr, n, 4, out.ptr, None)


def test_reduce_validator_shape(qo, orc):
    """validateReduce on generateRand data (values < TEST_MAX): SUM wraps."""
    n = 1 << 24
    r = gen_records(orc, n, 42, 1 << 30)
    s = int(r[:, 1].astype(np.uint64).sum() % (1 << 32))
    assert qo.operators.reduce(r, 0) == s
    assert qo.operators.reduce(r, 3) == s // n
    assert qo.operators.reduce(r, 1) == int(r[:, 1].max())
    assert qo.operators.reduce(r, 2) == int(r[:, 1].min())


# ---- point selection and device-buffer selection --------------------------------------------
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_point_selection(qo, orc, n):
    r = gen_records(orc, n, 50, 64)
    key = int(r[n // 2, 1])
    got = qo.operators.CL_PointSelection(r, key)
    assert np.array_equal(got, orc.filter_records(r, key, key))
    assert np.all(got[:, 1] == key)


def test_device_buffer_selection(qo, orc):
    """CL_PointSelectionOnly / CL_RangeSelectionOnly (Selection.cpp:92,111) over device buffers."""
    lib = qo._lib.load()
    n = 123_457
    r = gen_records(orc, n, 51, 100)
    with DeviceBuffer.from_numpy(r) as d:
        m = lib.qo_mem_wrap(d.ptr, d.nbytes)
        for fn, args, lo, hi in (("CL_PointSelectionOnly", (7,), 7, 7), ("CL_RangeSelectionOnly", (10, 20), 10, 20)):
            out = C.c_void_p()
            k = getattr(lib, fn)(m, n, *args, C.byref(out), 256, 128, 0)
            exp = orc.filter_records(r, lo, hi)
            assert k == exp.shape[0]
            host = np.empty((max(k, 1), 2), np.uint32)
            qo._lib.call("qo_mem_read", host.ctypes.data, out, k * 8)
            assert np.array_equal(host[:k], exp)
            lib.qo_mem_free(out)
        lib.qo_mem_free(m)


# ---- Record -> columns ----------------------------------------------------------------------
@pytest.mark.parametrize("n,off", [(1, 0), (1003, 0), (1003, 8), (1 << 22, 0)])
def test_unpack_records(qo, orc, n, off):
    r = np.stack([orc.gen_uniform_u32(n + 1, 60, 0, 0, 0), orc.gen_uniform_u32(n + 1, 60, 1, 0, 0)], 1)
    with DeviceBuffer.from_numpy(r) as d, DeviceBuffer(n * 4) as dx, DeviceBuffer(n * 4) as dy:
        qo._lib.call("qo_unpack_records", C.c_void_p(d.ptr + off), n, dx.ptr, dy.ptr, None)
        src = r[off // 8: off // 8 + n]
        assert np.array_equal(dx.to_numpy(np.uint32, n), src[:, 0])
        assert np.array_equal(dy.to_numpy(np.uint32, n), src[:, 1])
        # round trip through qo_pack_records
        with DeviceBuffer(n * 8) as back:
            qo._lib.call("qo_pack_records", dx.ptr, dy.ptr, n, back.ptr, None)
            assert np.array_equal(back.to_numpy(np.uint32, 2 * n).reshape(-1, 2), src)
