"""CPU tests that pin the oracle (oracle/qo_oracle.c) against hand-computed known
answers, brute-force fixtures, closed forms and invariants, so that a wrong
term, sign, index or operand in it fails here."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "known_answers.json").read_text())
BF = json.loads((GOLD / "brute_force.json").read_text())


def rec(a):
    return np.array(a, dtype=np.uint32).reshape(-1, 2)


def test_splitmix64_published_vector(orc):
    # splitmix64 from state 0: first output (Vigna's reference implementation).
    assert hex(orc.load().or_mix64(0x9E3779B97F4A7C15)) == KA["splitmix64_first_output_state0"]


def test_generator_properties(orc):
    u = orc.gen_uniform_u32(200_000, 7, 3, 0, 1000)
    assert u.min() >= 0 and u.max() < 1000
    hist = np.bincount(u, minlength=1000)
    assert hist.min() > 120 and hist.max() < 290  # ~200 expected per bucket
    for n in (1, 2, 3, 1000, 4096, 5000):
        p = orc.gen_perm(n, 99)
        assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint32))
    for bits in (5, 12, 31):
        xs = [orc.bij(i, bits, 5) for i in range(min(1 << bits, 4096))]
        assert len(set(xs)) == len(xs) and max(xs) < (1 << bits)
    t = orc.zipf_table(1000, 1.0)
    assert np.all(np.diff(t.astype(np.float64)) >= 0) and int(t[-1]) == 1 << 53
    keys = orc.gen_zipf_keys(100_000, t, 1, 2, 3)
    h = np.bincount(keys, minlength=1000)
    top = np.sort(h)[::-1]
    # Zipf z=1 over 1000 ranks: p(rank 0) = 1/H_1000 ~ 0.1336
    assert abs(top[0] / 100_000 - 0.1336) < 0.01
    assert keys.max() < 1000


def test_scan_known_answers(orc):
    for c in KA["scan"]:
        assert orc.scan_excl_i32(c["in"]).tolist() == c["out"]
    for c in BF["scan"]:
        assert orc.scan_excl_i32(c["in"]).tolist() == c["out"]


def test_scan_closed_forms(orc):
    n = 100_000
    assert np.array_equal(orc.scan_excl_i32(np.ones(n, np.int32)), np.arange(n, dtype=np.int32))
    x = np.full(5, 1 << 30, np.int32)
    assert orc.scan_excl_i32(x).tolist() == [0, 1 << 30, -(1 << 31), -(1 << 30), 0]


def test_filter_known_answers(orc):
    for c in KA["filter"] + BF["filter"]:
        out = orc.filter_records(rec(c["in"]), c["lo"], c["hi"])
        assert out.tolist() == (rec(c["out"]).tolist() if c["out"] else [])


def test_filter_agg_known_answers(orc):
    c = KA["filter_agg"]
    assert orc.filter_agg(c["key"], c["val"], c["lo"], c["hi"]) == (c["count"], c["sum"], c["sum32"])
    r = np.stack([np.array(c["val"], np.int32).view(np.uint32), np.array(c["key"], np.int32).view(np.uint32)], 1)
    assert orc.filter_agg_records(r, c["lo"], c["hi"]) == (c["count"], c["sum"], c["sum32"])


def test_filter_agg_matches_filter_then_sum(orc):
    rng = np.random.default_rng(1)
    k = rng.integers(-1000, 1000, 50_000, dtype=np.int32)
    v = rng.integers(-(1 << 31), 1 << 31, 50_000, dtype=np.int64).astype(np.int32)
    cnt, s, s32 = orc.filter_agg(k, v, -10, 300)
    sel = (k >= -10) & (k <= 300)
    assert cnt == int(sel.sum()) and s == int(v[sel].astype(np.int64).sum())
    assert s32 == int(np.int64(s).astype(np.int32))


def test_sort_known_answers(orc):
    for c in KA["sort"] + BF["sort"]:
        assert orc.sort_records(rec(c["in"])).tolist() == rec(c["out"]).tolist()


def test_sort_invariants(orc):
    rng = np.random.default_rng(2)
    r = np.stack([np.arange(30_000, dtype=np.uint32), rng.integers(0, 50, 30_000).astype(np.uint32)], 1)
    s = orc.sort_records(r)
    assert np.all(np.diff(s[:, 1].astype(np.int64)) >= 0)
    # stable: within equal keys the row ids (x) ascend
    same = s[1:, 1] == s[:-1, 1]
    assert np.all(s[1:, 0][same] > s[:-1, 0][same])
    assert sorted(map(tuple, s.tolist())) == sorted(map(tuple, r.tolist()))


def test_startpos_known_answers(orc):
    for c in KA["startpos"]:
        assert orc.group_startpos(rec(c["in"])).tolist() == c["out"]
    for c in BF["groupby_sorted"]:
        assert orc.group_startpos(rec(c["in"])).tolist() == c["startPos"]


def test_parallel_aggregate_quirks(orc):
    c = KA["parallel_aggregate"]
    for op, name in ((0, "sum"), (1, "max"), (2, "min"), (3, "avg")):
        assert orc.parallel_aggregate(rec(c["in"]), c["startPos"], op).tolist() == c[name], name
    for c in BF["groupby_sorted"]:
        for op in range(4):
            assert orc.parallel_aggregate(rec(c["in"]), c["startPos"], op).tolist() == c["agg"][str(op)]


def test_groupby_hash_known_answers(orc):
    c = KA["groupby"]
    got = np.stack([a.astype(np.int64) for a in orc.groupby_agg(c["key"], c["val"])], 1).tolist()
    assert got == c["out"]
    for c in BF["groupby_hash"]:
        got = orc.groupby_agg(c["key"], c["val"])
        got = np.stack([a.astype(np.int64) for a in got], 1).tolist() if len(got[0]) else []
        assert got == c["out"]


def test_join_known_answers(orc):
    c = KA["join"]
    rows = np.stack(orc.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"], c["sfilt"], -(1 << 31), 0), 1)
    assert rows.tolist() == c["filtered_fhi0"]
    rows = np.stack(orc.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"]), 1)
    assert rows.tolist() == c["unfiltered"]
    for c in BF["join"]:
        rows = orc.hash_join(c["rkey"], c["rpay"], c["skey"], c["spay"], c["sfilt"], c["flo"], c["fhi"])
        got = np.stack(rows, 1).tolist() if len(rows[0]) else []
        assert got == c["rows"]
        srt = orc.sort_join_rows(*rows)
        got = np.stack(srt, 1).tolist() if len(rows[0]) else []
        assert got == c["sorted"]


def test_join_closed_form_permutation(orc):
    # R keys a permutation of [0, nr): every selected S row matches exactly one R row.
    nr, ns = 5000, 40_000
    rk = orc.gen_perm(nr, 11).astype(np.int32)
    rp = np.arange(nr, dtype=np.int32) * 3
    sk = orc.gen_uniform_u32(ns, 1, 0, 0, nr).astype(np.int32)
    sp = np.arange(ns, dtype=np.int32)
    sf = orc.gen_uniform_u32(ns, 1, 1, 0, 100).astype(np.int32)
    k, r, s = orc.hash_join(rk, rp, sk, sp, sf, -(1 << 31), 49)
    sel = sf <= 49
    assert len(k) == int(sel.sum())
    inv = np.empty(nr, np.int64)
    inv[rk] = np.arange(nr)
    assert np.array_equal(s, sp[sel]) and np.array_equal(k, sk[sel]) and np.array_equal(r, rp[inv[sk[sel]]])


def test_join_sort_known_answers(orc):
    c = KA["join_sort"]
    a = np.array(c["in"], np.int32)
    got = np.stack(orc.sort_join_rows(a[:, 0], a[:, 1], a[:, 2]), 1).tolist()
    assert got == c["out"]


def test_probe_pairs_shape(orc):
    r = np.array([[5, 50], [7, 70], [5, 51]], np.uint32)
    s = np.array([[5, 1], [6, 2], [7, 3]], np.uint32)
    assert orc.probe_pairs(r, s).tolist() == [[1, 50], [1, 51], [3, 70]]


@pytest.mark.parametrize("mutant", ["drop_last", "wrong_sign"])
def test_planted_mistakes_are_caught(orc, mutant):
    """A deliberately broken restatement (computed here, on the CPU) must fail the
    same fixtures the oracle passes: the checks above have teeth."""
    c = KA["parallel_aggregate"]
    good = orc.parallel_aggregate(rec(c["in"]), c["startPos"], 0).tolist()
    assert good == c["sum"]
    if mutant == "drop_last":
        bad = [int(np.int64(sum(np.array(rec(c["in"])[s:e, 0], np.int64).astype(np.int32)[:-1])))
               for s, e in zip(c["startPos"], c["startPos"][1:] + [5])]
    else:
        bad = [-v for v in good]
    assert bad != c["sum"]


# ---- split / gather / scatter / projection / reduce ---------------------------------------
def _pid_ok(out, parts):
    """validateSplit (host/host/Validate.cpp:510-527): y % numPart never decreases."""
    pid = out[:, 1].astype(np.uint64) % parts
    return bool(np.all(np.diff(pid.astype(np.int64)) >= 0))


@pytest.mark.parametrize("i", range(6))
def test_split_brute_force(orc, i):
    c = BF["split"][i]
    out, bound = orc.split_records(rec(c["in"]), c["parts"])
    assert out.tolist() == c["out"]
    assert bound.tolist() == c["bound"]
    assert _pid_ok(out, c["parts"])


def test_split_invariants(orc):
    """Permutation of the input, validateSplit's order, stability (row ids rise
    inside a partition), and bounds = exclusive prefix of the partition counts."""
    n, parts = 20_000, 37
    r = np.stack([np.arange(n, dtype=np.uint32), orc.gen_uniform_u32(n, 5, 0, 0, 0)], 1)
    out, bound = orc.split_records(r, parts)
    assert _pid_ok(out, parts)
    assert np.array_equal(np.sort(out[:, 0]), r[:, 0])
    pid = out[:, 1] % parts
    for p in (0, 5, 36):
        ids = out[pid == p, 0]
        assert np.all(np.diff(ids.astype(np.int64)) > 0)
    counts = np.bincount(r[:, 1] % parts, minlength=parts)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert np.array_equal(bound[:, 0], starts) and np.array_equal(bound[:, 1], starts + counts)


@pytest.mark.parametrize("i", range(4))
def test_gather_brute_force(orc, i):
    c = BF["gather"][i]
    assert orc.gather_records(rec(c["R"]), c["loc"], rec(c["S0"])).tolist() == c["S"]


@pytest.mark.parametrize("i", range(4))
def test_scatter_brute_force(orc, i):
    c = BF["scatter"][i]
    assert orc.scatter_records(rec(c["R"]), c["loc"], rec(c["S0"])).tolist() == c["S"]


def test_gather_scatter_closed_forms(orc):
    """On a permutation p: gather(R, p) = R[p] (numpy fancy indexing), scatter
    inverts it, and scatter(gather(R, p), p) = R."""
    n = 5000
    r = np.stack([orc.gen_uniform_u32(n, 1, 0, 0, 0), orc.gen_uniform_u32(n, 1, 1, 0, 0)], 1)
    p = orc.gen_perm(n, 9).astype(np.int32)
    g = orc.gather_records(r, p)
    assert np.array_equal(g, r[p])
    assert np.array_equal(orc.scatter_records(g, p), r)
    s = np.zeros_like(r)
    s[p] = r
    assert np.array_equal(orc.scatter_records(r, p), s)


@pytest.mark.parametrize("i", range(3))
def test_project_brute_force(orc, i):
    c = BF["project"][i]
    assert orc.project_records(rec(c["R"]), rec(c["proj"])).tolist() == c["out"]


@pytest.mark.parametrize("i", range(5))
def test_reduce_brute_force(orc, i):
    c = BF["reduce"][i]
    assert [orc.reduce_records(rec(c["in"]), op) for op in range(4)] == c["out"]


def test_reduce_closed_forms(orc):
    """y = 0..n-1 (a permutation of it): SUM = n(n-1)/2 mod 2^32, MAX = n-1,
    MIN = 0, AVERAGE = floor(SUM mod 2^32 / n); and unsigned comparison: a y of
    2^31 is the maximum, not a negative minimum."""
    n = 100_003
    y = orc.gen_perm(n, 4)
    r = np.stack([np.zeros(n, np.uint32), y], 1)
    s = (n * (n - 1) // 2) % (1 << 32)
    assert [orc.reduce_records(r, op) for op in range(4)] == [s, n - 1, 0, s // n]
    q = rec([[0, 5], [0, 1 << 31], [0, 3]])
    assert orc.reduce_records(q, 1) == 1 << 31 and orc.reduce_records(q, 2) == 3


def test_unpack_records(orc):
    r = np.stack([orc.gen_uniform_u32(1001, 2, 0, 0, 0), orc.gen_uniform_u32(1001, 2, 1, 0, 0)], 1)
    x, y = orc.unpack_records(r)
    assert np.array_equal(x, r[:, 0]) and np.array_equal(y, r[:, 1])


def test_split_planted_mistake_is_caught(orc):
    """A split that orders partitions by y // parts (a transposed operand) or
    loses stability fails the fixtures the oracle passes."""
    c = BF["split"][3]
    recs = rec(c["in"])
    wrong_key = recs[np.argsort(recs[:, 1] // c["parts"], kind="stable")]
    assert wrong_key.tolist() != c["out"]
    unstable = recs[np.argsort(recs[:, 1] % c["parts"], kind="stable")][::-1]
    unstable = unstable[np.argsort(unstable[:, 1] % c["parts"], kind="stable")]
    assert unstable.tolist() != c["out"]
