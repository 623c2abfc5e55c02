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
