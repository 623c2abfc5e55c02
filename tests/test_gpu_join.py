"""GPU parity: hash build/probe with the fused S filter, hj_filter_sort, and the
reference-shaped bucket build/probe, against the oracle."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "known_answers.json").read_text())
BF = json.loads((GOLD / "brute_force.json").read_text())
INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1


def rows(cols):
    return np.stack(cols[:3], 1) if len(cols[0]) else np.zeros((0, 3), np.int32)


def canon(r):
    r = np.asarray(r, np.int64).reshape(-1, 3)
    return r[np.lexsort((r[:, 1], r[:, 2], r[:, 0]))] if len(r) else r


def test_join_known_answers(qo):
    c = KA["join"]
    got = qo.operators.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"], c["sfilt"], INT_MIN, 0)
    assert rows(got).tolist() == c["filtered_fhi0"] and got[3] == 1
    got = qo.operators.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"])
    # duplicate build key 5: matches of one S row come out in table-chain order
    assert np.array_equal(canon(rows(got)), canon(c["unfiltered"]))


def test_join_brute_force_fixtures(qo):
    for c in BF["join"]:
        got = qo.operators.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"], c["sfilt"], c["flo"], c["fhi"])
        assert np.array_equal(canon(rows(got)), canon(c["rows"]))
        srt = qo.operators.hj_filter_sort(c["rkey"], c["rpay"], c["skey"], c["spay"], c["sfilt"], c["flo"], c["fhi"])
        assert np.array_equal(canon(rows(srt)), canon(c["sorted"]))
        if len(c["rkey"]) == len(set(c["rkey"])):
            assert rows(srt).tolist() == c["sorted"]


@pytest.mark.parametrize("nr,ns", [(1, 1), (1000, 50_000), (1 << 16, 1 << 20), (1 << 20, (1 << 22) + 3)])
def test_join_unique_keys_matches_oracle_exactly(qo, orc, nr, ns):
    rk = orc.gen_perm(nr, 5).astype(np.int32)
    rp = orc.gen_uniform_u32(nr, 6, 0, 0, 0).astype(np.int32)
    sk = orc.gen_perm(nr, 5)[orc.gen_uniform_u32(ns, 6, 1, 0, nr)].astype(np.int32)
    sp = orc.gen_uniform_u32(ns, 6, 2, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, 6, 3, 0, 100).astype(np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, INT_MIN, 49)
    exp = orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, 49)
    assert got[3] == len(exp[0]) == int((sf <= 49).sum())
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_misses_and_no_filter(qo, orc):
    rng = np.random.default_rng(1)
    rk = rng.choice(np.arange(-10**6, 10**6), 20_000, replace=False).astype(np.int32)
    rp = rng.integers(INT_MIN, INT_MAX, 20_000, dtype=np.int64).astype(np.int32)
    sk = rng.integers(-10**6, 10**6, 300_000, dtype=np.int64).astype(np.int32)
    sp = np.arange(300_000, dtype=np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp)
    exp = orc.hash_join(rk, rp, sk, sp)
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_duplicates_and_sentinel(qo, orc):
    rng = np.random.default_rng(2)
    rk = rng.integers(-50, 50, 5000, dtype=np.int64).astype(np.int32)
    rp = rng.integers(-3, 3, 5000, dtype=np.int64).astype(np.int32)
    rk[:3] = -1
    rp[:2] = -1  # two (-1, -1) build tuples: the empty-slot bit pattern
    sk = rng.integers(-60, 60, 20_000, dtype=np.int64).astype(np.int32)
    sp = np.arange(20_000, dtype=np.int32)
    sf = rng.integers(0, 10, 20_000, dtype=np.int64).astype(np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, 2, 7, capacity=10_000_000)
    exp = orc.hash_join(rk, rp, sk, sp, sf, 2, 7)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_capacity_overflow_counts(qo, orc):
    rk = np.arange(100, dtype=np.int32)
    sk = np.tile(rk, 10)
    got = qo.operators.hash_join(rk, rk, sk, sk, capacity=50)
    assert got[3] == 1000 and len(got[0]) == 50
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[0], got[2])  # 50 genuine matches


def test_join_empty_sides(qo):
    e = np.zeros(0, np.int32)
    got = qo.operators.hash_join(e, e, np.arange(10, dtype=np.int32), np.arange(10, dtype=np.int32))
    assert got[3] == 0
    got = qo.operators.hash_join(np.arange(10, dtype=np.int32), np.arange(10, dtype=np.int32), e, e)
    assert got[3] == 0


@pytest.mark.parametrize("nr,mult", [(4096, 4), (100_000, 16)])
def test_hj_filter_sort_matches_oracle(qo, orc, nr, mult):
    ns = nr * mult
    rk = np.array([orc.bij(i, 31, 3) for i in range(nr)], np.uint32).astype(np.int32)
    rp = orc.gen_uniform_u32(nr, 2, 0, 0, 0).astype(np.int32)
    sk = rk[orc.gen_uniform_u32(ns, 2, 1, 0, nr)]
    sp = orc.gen_uniform_u32(ns, 2, 2, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, 2, 3, 0, 1000).astype(np.int32)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, INT_MIN, 199)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, 199))
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("base,nr", [(0, 1 << 16), (1000, 256), (-5000, 4096), (-100, 201), (INT_MAX - 4095, 4096),
                                     (INT_MIN, 4096), (77, 1)])
def test_hj_filter_sort_key_ranges(qo, orc, base, nr):
    """The sort skips radix passes over key bits that every build key shares;
    build keys in narrow ranges (positive, negative, across zero, at both ends
    of int32, a single key) must still come out fully ordered by (key, s_payload)."""
    rng = np.random.default_rng(nr)
    rk = (np.int64(base) + rng.permutation(nr)).astype(np.int32)
    rp = rng.integers(INT_MIN, INT_MAX, nr, dtype=np.int64).astype(np.int32)
    ns = max(64, nr * 8)
    sk = rk[rng.integers(0, nr, ns)]
    sp = rng.integers(INT_MIN, INT_MAX, ns, dtype=np.int64).astype(np.int32)
    sf = rng.integers(0, 1000, ns).astype(np.int32)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, INT_MIN, 599)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, 599))
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("longest", [64, 65])
def test_hj_filter_sort_run_lengths(qo, orc, longest):
    """Rows are sorted on the key bits first and every run of equal key by
    s_payload after, in tiles of 4096 positions; runs of up to 64 rows are
    sorted there, a longer one sends the whole output through the full-width
    sort.  Run lengths 1..longest (many runs crossing tile edges) and repeated
    s_payload values within runs."""
    rng = np.random.default_rng(longest)
    nr = 20_000
    rk = (np.int64(-7_000) + rng.permutation(nr)).astype(np.int32)
    rp = rng.integers(INT_MIN, INT_MAX, nr, dtype=np.int64).astype(np.int32)
    counts = rng.integers(0, 9, nr)
    counts[rng.choice(nr, 300, replace=False)] = longest
    counts[rng.choice(nr, 300, replace=False)] = longest - 1
    sk = np.repeat(rk, counts)
    sk = sk[rng.permutation(len(sk))]
    ns = len(sk)
    sp = rng.integers(-40, 40, ns).astype(np.int32)
    sf = np.zeros(ns, np.int32)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, INT_MIN, INT_MAX)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, INT_MAX))
    assert len(got[0]) == ns
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_hj_filter_sort_heavy_keys(qo, orc):
    """A third of the probe rows carry one key and the rest few keys: runs of
    ~10^5 rows take the full-width sort."""
    rng = np.random.default_rng(5)
    nr = 1000
    rk = rng.permutation(nr).astype(np.int32) * 3
    rp = rng.integers(INT_MIN, INT_MAX, nr, dtype=np.int64).astype(np.int32)
    ns = 600_000
    sk = rk[rng.integers(0, 40, ns)]
    sk[rng.random(ns) < 0.33] = rk[7]
    sp = rng.integers(INT_MIN, INT_MAX, ns, dtype=np.int64).astype(np.int32)
    sf = rng.integers(0, 1000, ns).astype(np.int32)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, INT_MIN, 499)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, 499))
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


# ---------------------------------------------------------------------------
# Table layouts and probe strategies:
#  * direct: hash-mixed table, direct build and probe;
#  * partitioned: hash-mixed table, partitioned (LDS-slice) build and the
#    one-split L2 probe, forced on small inputs with small buckets so skewed
#    keys exercise the overflow path;
#  * lds: hash-mixed table, partitioned build and the two-split LDS-slice
#    probe, with a 64 KiB table slice per partition (many partitions);
#  * keyrange: key-range table where the keys allow it, with the partitioned
#    paths planned (builds that fall back to hashing take the partitioned
#    build).
# ---------------------------------------------------------------------------
TUNE_JOIN_PART = 5
TUNE_JOIN_RANGE = 7


def _set_tune(qo, op, block, items, param):
    """Sets every size class of `op`; returns what each class read before.
    All classes are read first: a class without an entry reads its nearest
    class's entry, which must not be one set here."""
    import ctypes as C
    saved = []
    for cls in range(40):
        b, i, p = C.c_int(), C.c_int(), C.c_int()
        qo._lib.call("qo_tune_get", op, cls, C.byref(b), C.byref(i), C.byref(p))
        saved.append((op, cls, b.value, i.value, p.value))
    for cls in range(40):
        qo._lib.call("qo_tune_set", op, cls, block, items, param)
    return saved


def _restore_tune(qo, saved):
    lib = qo._lib.load()
    for op, cls, b, i, p in saved:
        lib.qo_tune_set(op, cls, b, i, p)


@pytest.fixture(params=["direct", "partitioned", "lds", "keyrange"])
def join_mode(request, qo):
    part, slice_kib, rng = {"direct": (2, 2048, 2), "partitioned": (1, 64, 2), "lds": (3, 64, 2),
                            "keyrange": (1, 64, 0)}[request.param]
    saved = _set_tune(qo, TUNE_JOIN_PART, 256, slice_kib, part) + _set_tune(qo, TUNE_JOIN_RANGE, 256, 256, rng)
    yield request.param
    _restore_tune(qo, saved)


def build_info(qo, rk, rp):
    """Builds a table from (rk, rp); returns (mode, slots in use, has_dups)."""
    import ctypes as C
    from query_on_opencl_fpga_amd.device import DeviceBuffer
    rk, rp = np.asarray(rk, np.int32), np.asarray(rp, np.int32)
    tb = qo._lib.load().qo_join_table_bytes(len(rk))
    with DeviceBuffer(tb) as t, DeviceBuffer.from_numpy(rk) as dk, DeviceBuffer.from_numpy(rp) as dp:
        qo._lib.call("qo_hash_build", t.ptr, tb, dk.ptr, dp.ptr, len(rk), None)
        mode, used, dups = C.c_int(), C.c_uint64(), C.c_int()
        qo._lib.call("qo_join_table_info", t.ptr, C.byref(mode), C.byref(used), C.byref(dups), None)
    return mode.value, used.value, dups.value


@pytest.mark.parametrize("base", [0, -12_345, INT_MIN, INT_MAX - 99_999, 1 << 30])
def test_key_range_table_matches_oracle(qo, orc, base):
    """Dense unique keys anywhere in the int32 range get the key-range table;
    probe keys just outside the range, and far outside it, miss."""
    nr, ns = 100_000, 1_000_003
    rk = (orc.gen_perm(nr, 41).astype(np.int64) + base).astype(np.int32)
    rp = orc.gen_uniform_u32(nr, 42, 0, 0, 0).astype(np.int32)
    sk = rk[orc.gen_uniform_u32(ns, 42, 1, 0, nr)]
    lo, hi = int(rk.min()), int(rk.max())
    sk[::11] = lo - 1 if lo > INT_MIN else INT_MAX
    sk[5::13] = hi + 1 if hi < INT_MAX else INT_MIN
    sk[7::17] = INT_MIN
    sk[9::19] = INT_MAX
    sp = orc.gen_uniform_u32(ns, 42, 2, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, 42, 3, 0, 100).astype(np.int32)
    mode, used, dups = build_info(qo, rk, rp)
    assert mode == 1 and used == 1 << 17 and dups == 0
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, 10, 79)
    exp = orc.hash_join(rk, rp, sk, sp, sf, 10, 79)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_key_range_falls_back_on_duplicates(qo, orc):
    """One duplicated key in a dense range: the build falls back to the
    hash-mixed table and every match of the duplicate is emitted."""
    nr, ns = 50_000, 400_000
    rk = orc.gen_perm(nr, 43).astype(np.int32)
    rk[123] = rk[4567]
    rp = np.arange(nr, dtype=np.int32)
    sk = orc.gen_perm(nr, 43)[orc.gen_uniform_u32(ns, 44, 1, 0, nr)].astype(np.int32)
    sk[::50] = rk[4567]
    sp = np.arange(ns, dtype=np.int32)
    assert build_info(qo, rk, rp)[0] == 0
    got = qo.operators.hash_join(rk, rp, sk, sp, capacity=2 * ns)
    exp = orc.hash_join(rk, rp, sk, sp)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))
    # the same key twice with the same payload is a duplicate too
    rk2, rp2 = np.arange(1000, dtype=np.int32), np.arange(1000, dtype=np.int32)
    rk2[10], rp2[10] = rk2[11], rp2[11]
    assert build_info(qo, rk2, rp2)[0] == 0
    got = qo.operators.hash_join(rk2, rp2, rk2, rp2, capacity=4000)
    exp = orc.hash_join(rk2, rp2, rk2, rp2)
    assert got[3] == len(exp[0]) == 1002
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_key_range_with_sentinel_build_tuples(qo, orc):
    """(-1, -1) build tuples are counted, not stored: a dense key set holding
    them (and another payload of key -1) keeps the key-range table and emits
    every match of key -1."""
    rk = np.arange(-1, 20_000, dtype=np.int32)
    rp = np.arange(rk.size, dtype=np.int32) * 7
    rp[0] = -1  # (-1, -1)
    rk = np.concatenate([rk, np.array([-1, -1], np.int32)])
    rp = np.concatenate([rp, np.array([-1, 55], np.int32)])  # a second (-1, -1) and (-1, 55)
    sk = np.concatenate([np.arange(-3, 20_003, dtype=np.int32), np.full(100, -1, np.int32)])
    sp = np.arange(sk.size, dtype=np.int32)
    mode, _, dups = build_info(qo, rk, rp)
    assert mode == 1 and dups == 1
    got = qo.operators.hash_join(rk, rp, sk, sp, capacity=100_000)
    exp = orc.hash_join(rk, rp, sk, sp)
    assert got[3] == len(exp[0]) == 20_000 + 101 * 3
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_key_range_not_chosen(qo, orc):
    """Keys spread wider than the table, or the key-range table switched off:
    the hash-mixed table."""
    rng = np.random.default_rng(45)
    rk = rng.choice(np.arange(-10**7, 10**7), 30_000, replace=False).astype(np.int32)
    assert build_info(qo, rk, rk)[0] == 0
    dense = np.arange(30_000, dtype=np.int32)
    assert build_info(qo, dense, dense)[0] == 1
    saved = _set_tune(qo, TUNE_JOIN_RANGE, 256, 256, 2)
    try:
        assert build_info(qo, dense, dense)[0] == 0
    finally:
        _restore_tune(qo, saved)


@pytest.mark.parametrize("nr,ns", [(1000, 3000), (50_000, 1_000_003), (1 << 20, 1 << 21)])
def test_join_modes_match_oracle(qo, orc, join_mode, nr, ns):
    rk = orc.gen_perm(nr, 15).astype(np.int32)
    rp = orc.gen_uniform_u32(nr, 16, 0, 0, 0).astype(np.int32)
    sk = orc.gen_perm(nr, 15)[orc.gen_uniform_u32(ns, 16, 1, 0, nr)].astype(np.int32)
    sk[::7] = -5  # misses
    sp = orc.gen_uniform_u32(ns, 16, 2, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, 16, 3, 0, 100).astype(np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, 10, 69)
    exp = orc.hash_join(rk, rp, sk, sp, sf, 10, 69)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))
    got = qo.operators.hash_join(rk, rp, sk, sp)  # no filter
    exp = orc.hash_join(rk, rp, sk, sp)
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_modes_skewed_probe_side(qo, orc, join_mode):
    """Most S tuples carry one key: its partition bucket overflows and the rest
    of its tuples are probed straight from the split pass."""
    nr, ns = 20_000, 400_000
    rk = orc.gen_perm(nr, 21).astype(np.int32)
    rp = np.arange(nr, dtype=np.int32) * 3
    sk = np.full(ns, rk[17], np.int32)
    sk[::5] = rk[orc.gen_uniform_u32(ns // 5, 22, 0, 0, nr)]
    sp = np.arange(ns, dtype=np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp)
    exp = orc.hash_join(rk, rp, sk, sp)
    assert got[3] == ns
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_modes_skewed_build_side(qo, orc, join_mode):
    """Thousands of build tuples share one key (and some are the (-1, -1)
    bit pattern): the build bucket overflows into direct inserts and the probe
    takes the duplicate-key path."""
    rng = np.random.default_rng(23)
    nr = 30_000
    rk = rng.permutation(np.arange(10, 10 + nr)).astype(np.int32)
    rk[:6000] = 77
    rk[6000:6003] = -1
    rp = rng.integers(-1000, 1000, nr, dtype=np.int64).astype(np.int32)
    rp[6000:6002] = -1
    sk = rng.integers(0, 12 + nr, 50_000, dtype=np.int64).astype(np.int32)
    sk[:40] = 77
    sk[40:50] = -1
    sp = np.arange(50_000, dtype=np.int32)
    sf = rng.integers(0, 4, 50_000, dtype=np.int64).astype(np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, 0, 2, capacity=1_000_000)
    exp = orc.hash_join(rk, rp, sk, sp, sf, 0, 2)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(rows(got)), canon(rows(exp)))


def test_join_modes_capacity_overflow(qo, join_mode):
    rk = np.arange(5000, dtype=np.int32)
    sk = np.tile(rk, 20)
    got = qo.operators.hash_join(rk, rk, sk, sk, capacity=777)
    assert got[3] == 100_000 and len(got[0]) == 777
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[0], got[2])


def test_hj_filter_sort_modes(qo, orc, join_mode):
    nr, ns = 60_000, 700_000
    rk = orc.gen_perm(nr, 31).astype(np.int32) - 30_000
    rp = orc.gen_uniform_u32(nr, 32, 0, 0, 0).astype(np.int32)
    sk = rk[orc.gen_uniform_u32(ns, 32, 1, 0, nr)]
    sp = orc.gen_uniform_u32(ns, 32, 2, 0, 1000).astype(np.int32)  # repeated (key, spay) ties
    sf = orc.gen_uniform_u32(ns, 32, 3, 0, 1000).astype(np.int32)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, 100, 599)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, 100, 599))
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


PART_L2 = 4  # tune operator 5: key-range tables probed through L2 only


def _lds_route_case(orc, nr, ns, seed, skew=False, sentinel=False):
    rk = (orc.gen_perm(nr, seed).astype(np.int64) - 777).astype(np.int32)
    rp = orc.gen_uniform_u32(nr, seed + 1, 0, 0, 0).astype(np.int32)
    # probe keys over the key range widened by 1000 at both ends (misses)
    sk = (orc.gen_uniform_u32(ns, seed + 1, 1, 0, nr + 2000).astype(np.int64) - 1777).astype(np.int32)
    if skew:  # 3 of 4 probe tuples carry one key: its bucket's regions overflow
        sk[np.arange(ns) % 4 != 0] = rk[5]
    if sentinel:  # a (-1, -1) build tuple: the table keeps the key-range mode, the probe takes the L2 route
        rp[np.flatnonzero(rk == -1)[0]] = -1  # key -1 is in the permutation's range
        sk[::101] = -1
    sp = orc.gen_uniform_u32(ns, seed + 1, 2, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, seed + 1, 3, 0, 100).astype(np.int32)
    return rk, rp, sk, sp, sf


@pytest.mark.parametrize("nr,ns,skew,sentinel", [
    (1 << 21, 1 << 22, False, False),  # capacity 2^21 in slots for 2^22: buckets take two planned regions
    (3 << 20, 1 << 22, False, False),  # capacity 2^22, as planned
    (3 << 20, 1 << 22, True, False),
    (3 << 20, 1 << 22, False, True),
    (1 << 24, 1 << 24, False, False),  # capacity 2^24: 512 buckets
])
def test_key_range_lds_route_matches_oracle(qo, orc, nr, ns, skew, sentinel):
    """Key-range tables of 2^21 .. 2^24 slots are probed through LDS slices of
    2^15 slots by default (a one-pass split by key into up to 512 buckets):
    rows equal to the oracle's, and to the L2 route's (tune operator 5,
    param 4).  Skewed probe keys overflow their bucket's regions into direct
    probes; a (-1, -1) build tuple sends the table to the L2 route."""
    rk, rp, sk, sp, sf = _lds_route_case(orc, nr, ns, 61, skew, sentinel)
    assert build_info(qo, rk, rp)[0] == 1  # key-range table
    exp = canon(rows(orc.hash_join(rk, rp, sk, sp, sf, 0, 49)))
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, 0, 49, capacity=ns + 16)
    assert got[3] == len(exp)
    assert np.array_equal(canon(rows(got)), exp)
    saved = _set_tune(qo, TUNE_JOIN_PART, 256, 32, PART_L2)
    try:
        got = qo.operators.hash_join(rk, rp, sk, sp, sf, 0, 49, capacity=ns + 16)
    finally:
        _restore_tune(qo, saved)
    assert got[3] == len(exp)
    assert np.array_equal(canon(rows(got)), exp)


@pytest.mark.parametrize("flo,fhi", [(100, 200), (0, 99), (50, 49)])
def test_key_range_lds_route_filter_edges(qo, orc, flo, fhi):
    """The LDS route with every probe tuple filtered out (each bucket empty:
    the probe's work-groups return before any barrier), with none filtered out
    (the fullest buckets), and with an empty filter range."""
    rk, rp, sk, sp, sf = _lds_route_case(orc, 3 << 20, 1 << 21, 81)
    exp = canon(rows(orc.hash_join(rk, rp, sk, sp, sf, flo, fhi)))
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, flo, fhi, capacity=(1 << 21) + 16)
    assert got[3] == len(exp)
    assert np.array_equal(canon(rows(got)), exp)


def test_hj_filter_sort_lds_route(qo, orc):
    """hj_filter_sort on a key-range table of 2^22 slots (the probe's
    sort-ready output through the LDS route): rows in (key, s_payload) order,
    equal to the oracle's."""
    rk, rp, sk, sp, sf = _lds_route_case(orc, 3 << 20, 1 << 23, 71)
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, 0, 49)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, 0, 49))
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_bucket_build_probe_matches_oracle(qo, orc):
    rng = np.random.default_rng(3)
    nr, ns, B, bs = 20_000, 100_000, 1 << 15, 8
    r = np.stack([rng.permutation(nr), rng.integers(0, 1 << 31, nr)], 1).astype(np.uint32)
    s = np.stack([rng.integers(0, 2 * nr, ns), np.arange(ns)], 1).astype(np.uint32)
    table, counts = qo.operators.buildHashTable(r, B, bs)
    got = qo.operators.probeHashTable(table, counts, s, B, bs)
    table.free()
    counts.free()
    exp = orc.probe_pairs(r, s)
    assert np.array_equal(got, exp)  # unique R keys: one match per hit, S order


def test_bucket_overflow_is_an_error(qo):
    r = np.stack([np.zeros(100), np.arange(100)], 1).astype(np.uint32)  # 100 tuples in one bucket
    with pytest.raises(qo._lib.QoError):
        qo.operators.buildHashTable(r, 16, 8)
