"""CPU tests of the multi-threaded CPU baseline (oracle/qo_cpu_par.c): on the same
seeded inputs it must return exactly what the single-threaded oracle returns
(which tests/test_oracle.py pins), for several thread counts, including more
threads than rows and inputs with duplicate keys."""
import ctypes as C

import numpy as np
import pytest


def _p(a):
    return a.ctypes.data


@pytest.mark.parametrize("threads", [1, 3, 8, 64])
@pytest.mark.parametrize("n", [0, 1, 7, 1000, 100_003])
def test_filter_agg_matches_oracle(orc, threads, n):
    k = orc.gen_uniform_u32(n, 11, 0, 0, 1 << 31).astype(np.int32)
    v = orc.gen_uniform_u32(n, 11, 1, 0, 0).astype(np.int32)  # full 32-bit values: the int32 sum wraps
    out = np.zeros(3, np.int64)
    orc.load().pr_filter_agg(_p(k), _p(v), n, 0, (1 << 30) - 1, _p(out), threads)
    assert tuple(out) == tuple(orc.filter_agg(k, v, 0, (1 << 30) - 1))


@pytest.mark.parametrize("threads", [1, 5, 16])
@pytest.mark.parametrize("n,G", [(1, 1), (50_000, 16), (200_000, 1024), (300_000, 100_000)])
def test_groupby_matches_oracle(orc, threads, n, G):
    t = orc.zipf_table(G, 0.8)
    k = orc.gen_zipf_keys(n, t, 3, 2, 5).astype(np.int32)
    k[::7] = -k[::7] - 1  # signed keys order below the others
    v = orc.gen_uniform_u32(n, 3, 3, 0, 1 << 16).astype(np.int32)
    exp = orc.groupby_agg(k, v)
    g = len(exp[0])
    outs = [np.zeros(n, d) for d in (np.int32, np.int64, np.int64, np.int32, np.int32)]
    got = orc.load().pr_groupby_agg(_p(k), _p(v), n, *[_p(o) for o in outs], threads)
    assert got == g
    for o, e in zip(outs, exp):
        assert np.array_equal(o[:g], e)


@pytest.mark.parametrize("threads", [1, 4, 32])
@pytest.mark.parametrize("dups", [False, True])
def test_hash_join_matches_oracle(orc, threads, dups):
    nr, ns = 20_000, 150_000
    rk = orc.gen_perm(nr, 3).astype(np.int32)
    if dups:
        rk[::5] = rk[1::5][: len(rk[::5])]  # repeated build keys: several matches per probe row
    rp = orc.gen_uniform_u32(nr, 3, 1, 0, 0).astype(np.int32)
    sk = orc.gen_uniform_u32(ns, 3, 2, 0, nr + 500).astype(np.int32)  # some fk miss
    sp = orc.gen_uniform_u32(ns, 3, 3, 0, 0).astype(np.int32)
    sf = orc.gen_uniform_u32(ns, 3, 4, 0, 100).astype(np.int32)
    exp = orc.hash_join(rk, rp, sk, sp, sf, -(1 << 31), 49)
    m = len(exp[0])
    outs = [np.zeros(max(m, 1), np.int32) for _ in range(3)]
    lib = orc.load()
    got = lib.pr_hash_join(_p(rk), _p(rp), nr, _p(sk), _p(sp), _p(sf), ns, -(1 << 31), 49,
                           *[_p(o) for o in outs], m, threads)
    assert got == m
    # Row order differs (per-thread ranges, chain order): compare sorted by (key, spay, rpay).
    lib.pr_sort_join_rows(_p(outs[0]), _p(outs[1]), _p(outs[2]), m, threads)
    ek, er, es = orc.sort_join_rows(*exp)
    o = np.lexsort((outs[1][:m], outs[2][:m], outs[0][:m]))
    e = np.lexsort((er, es, ek))
    assert np.array_equal(outs[0][:m][o], ek[e])
    assert np.array_equal(outs[1][:m][o], er[e])
    assert np.array_equal(outs[2][:m][o], es[e])


@pytest.mark.parametrize("threads", [1, 6])
def test_sort_join_rows_matches_oracle(orc, threads):
    m = 123_457
    k = orc.gen_uniform_u32(m, 9, 0, 0, 5000).astype(np.int32) - 2500
    r = orc.gen_uniform_u32(m, 9, 1, 0, 0).astype(np.int32)
    s = orc.gen_uniform_u32(m, 9, 2, 0, 64).astype(np.int32) - 32  # ties on (key, spay)
    a = [x.copy() for x in (k, r, s)]
    b = [x.copy() for x in (k, r, s)]
    orc.load().pr_sort_join_rows(_p(a[0]), _p(a[1]), _p(a[2]), m, threads)
    b = orc.sort_join_rows(*b)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)  # both stable: identical, ties included
