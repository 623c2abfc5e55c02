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


# ---------------------------------------------------------------------------
# Direct vs partitioned build/probe (the partitioned path is forced on small
# inputs with a 64 KiB table slice per partition, i.e. many partitions, and
# small buckets so skewed keys exercise the overflow path).
# ---------------------------------------------------------------------------
TUNE_JOIN_PART = 5


@pytest.fixture(params=["direct", "partitioned"])
def join_mode(request, qo):
    import ctypes as C
    lib = qo._lib.load()
    mode, slice_kib = (2, 2048) if request.param == "direct" else (1, 64)
    saved = []
    for cls in range(40):
        b, i, p = C.c_int(), C.c_int(), C.c_int()
        qo._lib.call("qo_tune_get", TUNE_JOIN_PART, cls, C.byref(b), C.byref(i), C.byref(p))
        saved.append((cls, b.value, i.value, p.value))
        qo._lib.call("qo_tune_set", TUNE_JOIN_PART, cls, 256, slice_kib, mode)
    yield request.param
    for cls, b, i, p in saved:
        lib.qo_tune_set(TUNE_JOIN_PART, cls, b, i, p)


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
