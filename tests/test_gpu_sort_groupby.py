"""GPU parity: Record sort, join-row sort, hash group-by aggregation and the
reference's sort-based group-by pipeline, against the oracle (bit-exact)."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "known_answers.json").read_text())
BF = json.loads((GOLD / "brute_force.json").read_text())
INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1


def test_sort_known_answers(qo):
    for c in KA["sort"] + BF["sort"]:
        assert qo.operators.sort_records(np.array(c["in"], np.uint32)).tolist() == c["out"]


@pytest.mark.parametrize("n", [1, 2, 3, 4095, 4096, 4097, 100_000, 1 << 20])
@pytest.mark.parametrize("kmax", [1, 50, 1 << 32])
def test_sort_records_matches_oracle(qo, orc, n, kmax):
    rng = np.random.default_rng(n + kmax)
    r = np.stack([np.arange(n), rng.integers(0, kmax, n, dtype=np.uint64)], 1).astype(np.uint32)
    assert np.array_equal(qo.operators.sort_records(r), orc.sort_records(r))


def test_sort_join_rows_matches_oracle(qo, orc):
    c = KA["join_sort"]
    a = np.array(c["in"], np.int32)
    got = np.stack(qo.operators.sort_join_rows(a[:, 0], a[:, 1], a[:, 2]), 1).tolist()
    assert got == c["out"]
    rng = np.random.default_rng(3)
    for m, kr in ((2, 5), (10_000, 100), (300_000, 1 << 31), (1 << 20, 1000)):
        k = rng.integers(-kr, kr, m, dtype=np.int64).astype(np.int32)
        r = rng.integers(INT_MIN, INT_MAX, m, dtype=np.int64).astype(np.int32)
        s = rng.integers(-50, 50, m, dtype=np.int64).astype(np.int32)
        got = qo.operators.sort_join_rows(k, r, s)
        exp = orc.sort_join_rows(k, r, s)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)


def _gb_rows(res):
    return np.stack([a.astype(np.int64) for a in res], 1).tolist() if len(res[0]) else []


def test_groupby_known_answers(qo):
    c = KA["groupby"]
    assert _gb_rows(qo.operators.groupby_agg(c["key"], c["val"], 8)) == c["out"]
    for c in BF["groupby_hash"]:
        assert _gb_rows(qo.operators.groupby_agg(c["key"], c["val"], max(1, len(c["key"])))) == c["out"]


@pytest.mark.parametrize("G,z", [(1, 0.0), (1024, 0.0), (1024, 1.0), (50_000, 0.5), (1 << 20, 0.0), (1 << 20, 1.0)])
def test_groupby_matches_oracle(qo, orc, G, z):
    n = 2_000_003
    t = orc.zipf_table(G, z)
    keys = orc.gen_zipf_keys(n, t, 9, 1, 17).astype(np.int32)
    vals = orc.gen_uniform_u32(n, 9, 2, 0, 1 << 16).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, G)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_signed_keys_and_sentinel(qo, orc):
    rng = np.random.default_rng(4)
    n = 500_000
    keys = rng.integers(-3000, 3000, n, dtype=np.int64).astype(np.int32)
    keys[::7] = -1  # the empty-slot bit pattern is a valid key
    vals = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, 6001)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_capacity_overflow_reports(qo):
    keys = np.arange(10_000, dtype=np.int32)
    with pytest.raises(qo._lib.QoError):
        qo.operators.groupby_agg(keys, keys, 100)


# ---------------------------------------------------------------------------
# Direct (LDS cache + global table) vs partitioned (hash split, per-bucket LDS
# tables) group-by; the partitioned path is forced on small inputs with 16
# groups per bucket so that both split levels and the overflow list are used.
# ---------------------------------------------------------------------------
TUNE_GROUPBY_PART = 6


@pytest.fixture(params=["direct", "partitioned"])
def gb_mode(request, qo):
    import ctypes as C
    lib = qo._lib.load()
    mode, target = (2, 1024) if request.param == "direct" else (1, 16)
    saved = []
    for cls in range(40):  # all read before any is set (a class without an entry reads its nearest)
        b, i, p = C.c_int(), C.c_int(), C.c_int()
        qo._lib.call("qo_tune_get", TUNE_GROUPBY_PART, cls, C.byref(b), C.byref(i), C.byref(p))
        saved.append((cls, b.value, i.value, p.value))
    for cls in range(40):
        qo._lib.call("qo_tune_set", TUNE_GROUPBY_PART, cls, 512, target, mode)
    yield request.param
    for cls, b, i, p in saved:
        lib.qo_tune_set(TUNE_GROUPBY_PART, cls, b, i, p)


@pytest.mark.parametrize("G,z,n", [(100, 0.0, 1000), (5000, 1.0, 300_000), (70_000, 0.5, 1_000_003),
                                   (1 << 18, 0.0, 2_000_000)])
def test_groupby_modes_match_oracle(qo, orc, gb_mode, G, z, n):
    t = orc.zipf_table(G, z)
    keys = orc.gen_zipf_keys(n, t, 11, 1, 29).astype(np.int32)
    vals = orc.gen_uniform_u32(n, 11, 2, 0, 1 << 16).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, G)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_modes_signed_keys_and_sentinel(qo, orc, gb_mode):
    rng = np.random.default_rng(14)
    n = 400_000
    keys = rng.integers(-20_000, 20_000, n, dtype=np.int64).astype(np.int32)
    keys[::9] = -1
    keys[5::11] = INT_MIN
    vals = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, 40_001)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_modes_heavy_hitter(qo, orc, gb_mode):
    """60% of the tuples carry one key: its bucket overflows and the excess is
    aggregated from the overflow list."""
    rng = np.random.default_rng(15)
    n = 1_500_000
    keys = rng.integers(0, 30_000, n, dtype=np.int64).astype(np.int32)
    keys[rng.random(n) < 0.6] = 12_345
    vals = rng.integers(-(1 << 20), 1 << 20, n, dtype=np.int64).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, 30_000)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_modes_heavy_sentinel_keys(qo, orc, gb_mode):
    """Heavy keys that are the empty-slot pattern (-1) and INT_MIN, plus a third
    heavy key: a bucket's dominant key is summed in registers and merged once,
    also when it is the key that never enters an LDS table."""
    rng = np.random.default_rng(16)
    n = 2_000_000
    keys = rng.integers(-50_000, 50_000, n, dtype=np.int64).astype(np.int32)
    u = rng.random(n)
    keys[u < 0.4] = -1
    keys[(u >= 0.4) & (u < 0.5)] = INT_MIN
    keys[(u >= 0.5) & (u < 0.6)] = 7
    vals = rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, 100_003)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_groupby_modes_capacity_overflow_reports(qo, gb_mode):
    keys = np.arange(20_000, dtype=np.int32)
    with pytest.raises(qo._lib.QoError):
        qo.operators.groupby_agg(keys, keys, 100)


def test_startpos_and_aggregate_known_answers(qo):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    c = KA["parallel_aggregate"]
    r = np.array(c["in"], np.uint32)
    sp = np.array(c["startPos"], np.int32)
    for op, name in ((0, "sum"), (1, "max"), (2, "min"), (3, "avg")):
        with DeviceBuffer.from_numpy(r) as dr, DeviceBuffer.from_numpy(sp) as dsp, DeviceBuffer(12) as out:
            qo._lib.call("qo_aggregate_groups", dr.ptr, r.shape[0], dsp.ptr, 3, op, out.ptr, None)
            assert out.to_numpy(np.int32, 3).tolist() == c[name], name


@pytest.mark.parametrize("n,kmax", [(1, 1), (1000, 3), (300_000, 50), (1 << 20, 1 << 20)])
def test_sort_based_groupby_matches_oracle(qo, orc, n, kmax):
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    rng = np.random.default_rng(kmax)
    r = np.stack([rng.integers(INT_MIN, INT_MAX, n, dtype=np.int64), rng.integers(0, kmax, n)], 1).astype(np.int64)
    r = r.astype(np.uint32)
    s = orc.sort_records(r)
    exp_sp = orc.group_startpos(s)
    with DeviceBuffer.from_numpy(s) as ds, DeviceBuffer(n * 4) as dsp, DeviceBuffer(8) as num:
        qo._lib.call("qo_group_startpos", ds.ptr, n, dsp.ptr, num.ptr, None)
        g = int(num.to_numpy(np.uint64, 1)[0])
        sp = dsp.to_numpy(np.int32, g)
        assert np.array_equal(sp, exp_sp)
        for op in range(4):
            with DeviceBuffer(max(g, 1) * 4) as out:
                qo._lib.call("qo_aggregate_groups", ds.ptr, n, dsp.ptr, g, op, out.ptr, None)
                assert np.array_equal(out.to_numpy(np.int32, g), orc.parallel_aggregate(s, sp, op)), op
