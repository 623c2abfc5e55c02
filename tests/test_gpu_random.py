"""GPU: a seeded randomized sweep of the operators against the oracle.

Each case draws a size around the kernels' tile and wave boundaries (or a
random one), a key distribution (full 32-bit range, a narrow dense range placed
anywhere in int32 including its ends, one heavy key, values around -1, already
sorted or reversed) and random predicate bounds, then compares the operator's
output with the oracle's: bit-exact, in the order the operator defines (the
join's rows as a multiset).  The fixed-shape tests elsewhere pin the
documented cases; this sweep looks for the combinations they miss."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1
EDGE = [1, 2, 3, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 4095, 4096, 4097, 8191, 8192, 8193, 16383, 16384,
        16385, 32767, 32769, 65535, 65537, 131071, 262145]
KINDS = ["uniform", "dense", "heavy", "around_minus_one", "sorted", "reversed"]


def size(rng):
    return int(rng.choice(EDGE)) if rng.random() < 0.7 else int(rng.integers(1, 300_000))


def keys(rng, n, kind):
    if kind == "uniform":
        k = rng.integers(INT_MIN, INT_MAX, n, endpoint=True)
    elif kind == "dense":
        span = int(rng.integers(1, 5000))
        base = int(rng.choice([0, -5000, 123_456, INT_MIN, INT_MAX - span + 1]))
        k = base + rng.integers(0, span, n)
    elif kind == "heavy":
        k = rng.integers(-1000, 1000, n)
        k[rng.random(n) < 0.9] = 7
    elif kind == "around_minus_one":
        k = rng.integers(-3, 3, n)
    elif kind == "sorted":
        k = np.sort(rng.integers(INT_MIN, INT_MAX, n, endpoint=True))
    else:
        k = np.sort(rng.integers(-100_000, 100_000, n))[::-1]
    return np.asarray(k, np.int64).astype(np.int32)


def bounds(rng):
    c = rng.integers(0, 5)
    if c == 0:
        return INT_MIN, INT_MAX
    if c == 1:
        return 10, 5  # empty
    a, b = sorted(int(x) for x in rng.integers(INT_MIN, INT_MAX, 2, endpoint=True))
    if c == 2:
        return -1000, 1000
    return a, b


def records(rng, n, kind):
    return np.stack([np.arange(n, dtype=np.uint32), keys(rng, n, kind).view(np.uint32)], 1)


def canon(r):
    r = np.asarray(r, np.int64).reshape(-1, 3)
    return r[np.lexsort((r[:, 1], r[:, 2], r[:, 0]))] if len(r) else r


@pytest.mark.parametrize("seed", range(16))
def test_random_filter_records(qo, orc, seed):
    rng = np.random.default_rng(1000 + seed)
    r = records(rng, size(rng), KINDS[seed % len(KINDS)])
    lo, hi = bounds(rng)
    assert np.array_equal(qo.operators.filter_records(r, lo, hi), orc.filter_records(r, lo, hi))


@pytest.mark.parametrize("seed", range(12))
def test_random_filter_agg(qo, orc, seed):
    rng = np.random.default_rng(2000 + seed)
    n = size(rng)
    k = keys(rng, n, KINDS[seed % len(KINDS)])
    v = keys(rng, n, "uniform")
    lo, hi = bounds(rng)
    assert qo.operators.filter_agg(k, v, lo, hi) == orc.filter_agg(k, v, lo, hi)


@pytest.mark.parametrize("seed", range(12))
def test_random_scan(qo, orc, seed):
    rng = np.random.default_rng(3000 + seed)
    x = keys(rng, size(rng), KINDS[seed % len(KINDS)])
    assert np.array_equal(qo.operators.scanImpl(x), orc.scan_excl_i32(x))


@pytest.mark.parametrize("seed", range(12))
def test_random_sort_records(qo, orc, seed):
    rng = np.random.default_rng(4000 + seed)
    r = records(rng, size(rng), KINDS[seed % len(KINDS)])
    assert np.array_equal(qo.operators.sort_records(r), orc.sort_records(r))


@pytest.mark.parametrize("seed", range(16))
def test_random_groupby(qo, orc, seed):
    rng = np.random.default_rng(5000 + seed)
    n = size(rng)
    k = keys(rng, n, KINDS[seed % len(KINDS)])
    v = keys(rng, n, "uniform")
    got = qo.operators.groupby_agg(k, v, n)
    exp = orc.groupby_agg(k, v)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)


def join_inputs(rng, seed, max_nr=None):
    nr = size(rng) if rng.random() < 0.5 else int(rng.integers(1, 20_000))
    if max_nr:
        nr = min(nr, max_nr)
    ns = int(rng.integers(1, 4 * nr + 2))
    kind = seed % 4
    if kind == 0:  # unique keys, a permutation from a random base (key-range table)
        base = int(rng.choice([0, -7, INT_MIN, INT_MAX - nr + 1, 1 << 30]))
        rk = (base + rng.permutation(nr)).astype(np.int64).astype(np.int32)
    elif kind == 1:  # unique sparse keys (hash-mixed table)
        rk = rng.choice(np.arange(INT_MIN, INT_MAX, 4099, dtype=np.int64), nr, replace=False).astype(np.int32)
    elif kind == 2:  # duplicate keys, -1 among them
        rk = rng.integers(-3, max(4, nr // 3), nr).astype(np.int32)
    else:  # unique keys with the (-1, -1) tuple
        rk = (rng.permutation(nr) - nr // 2).astype(np.int32)
    rp = keys(rng, nr, "uniform")
    if kind == 3 and (rk == -1).any():
        rp[rk == -1] = -1
    hit = rk[rng.integers(0, nr, ns)]
    miss = keys(rng, ns, "uniform")
    sk = np.where(rng.random(ns) < 0.8, hit, miss).astype(np.int32)
    sp = keys(rng, ns, "uniform")
    sf = rng.integers(0, 100, ns).astype(np.int32)
    flo, fhi = (INT_MIN, int(rng.integers(0, 100))) if rng.random() < 0.8 else (INT_MIN, INT_MAX)
    return rk, rp, sk, sp, sf, flo, fhi


@pytest.mark.parametrize("seed", range(16))
def test_random_join(qo, orc, seed):
    rng = np.random.default_rng(6000 + seed)
    rk, rp, sk, sp, sf, flo, fhi = join_inputs(rng, seed)
    exp = orc.hash_join(rk, rp, sk, sp, sf, flo, fhi)
    # Duplicate build keys can give more matches than S has tuples.
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, flo, fhi, capacity=max(len(exp[0]), 1))
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(np.stack(got[:3], 1)), canon(np.stack(exp, 1)))


@pytest.mark.parametrize("seed", range(12))
def test_random_hj_filter_sort(qo, orc, seed):
    rng = np.random.default_rng(7000 + seed)
    rk, rp, sk, sp, sf, flo, fhi = join_inputs(rng, seed)
    exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, flo, fhi))
    got = qo.operators.hj_filter_sort(rk, rp, sk, sp, sf, flo, fhi, capacity=max(len(exp[0]), 1))
    # Exact in (key, s_payload); rows equal in both (an S tuple matching a
    # duplicate build key several times) come in no specified order.
    for a, b in zip(got[:1] + got[2:], exp[:1] + exp[2:]):
        assert np.array_equal(a, b)
    assert np.array_equal(canon(np.stack(got, 1)), canon(np.stack(exp, 1)))


@pytest.mark.parametrize("seed", range(12))
def test_random_split(qo, orc, seed):
    rng = np.random.default_rng(8000 + seed)
    r = records(rng, size(rng), KINDS[seed % len(KINDS)])
    parts = int(rng.choice([1, 2, 3, 255, 256, 257, 1000, 4096, 65536]))
    out, bound = qo.operators.splitImpl(r, parts)
    eo, eb = orc.split_records(r, parts)
    assert np.array_equal(out, eo) and np.array_equal(bound, eb)


@pytest.mark.parametrize("seed", range(12))
def test_random_reduce(qo, orc, seed):
    rng = np.random.default_rng(9000 + seed)
    r = records(rng, size(rng), KINDS[seed % len(KINDS)])
    op = seed % 4
    assert qo.operators.reduce(r, op) == orc.reduce_records(r, op)


@pytest.mark.parametrize("seed", range(6))
def test_random_batch(qo, orc, seed):
    """Batches of 8 random queries of every key kind (one graph launch), each
    query against the oracle with the same tie rule as the plain path."""
    rng = np.random.default_rng(10_000 + seed)
    qs = [join_inputs(rng, seed * 8 + i, max_nr=20_000) for i in range(8)]
    exps = [orc.sort_join_rows(*orc.hash_join(*q)) for q in qs]
    got = qo.operators.batch_hj_filter_sort(qs, capacities=[max(len(e[0]), 1) for e in exps])
    for g, e in zip(got, exps):
        for a, b in zip(g[:1] + g[2:], e[:1] + e[2:]):
            assert np.array_equal(a, b)
        assert np.array_equal(canon(np.stack(g, 1)), canon(np.stack(e, 1)))


@pytest.mark.parametrize("seed", range(12))
def test_random_groupby_filtered(qo, orc, seed):
    """The fused filter -> group-by, with the filter on its own column or on the
    group key itself, against the oracle's filter then group-by."""
    from test_gpu_groupby_filtered import run_filtered

    rng = np.random.default_rng(11_000 + seed)
    n = size(rng)
    k = keys(rng, n, KINDS[seed % len(KINDS)])
    v = keys(rng, n, "uniform")
    f = k if seed % 2 else keys(rng, n, "dense")
    lo, hi = bounds(rng)
    got = run_filtered(qo, k, v, f, lo, hi, n)
    sel = (f >= lo) & (f <= hi)
    exp = orc.groupby_agg(k[sel], v[sel])
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("seed", range(12))
def test_random_gather_scatter_projection(qo, orc, seed):
    """Locations and row ids anywhere in int32, most of them in range."""
    rng = np.random.default_rng(12_000 + seed)
    r = records(rng, size(rng), KINDS[seed % len(KINDS)])
    n = r.shape[0]
    m = size(rng)
    loc = np.where(rng.random(m) < 0.9, rng.integers(0, n, m), keys(rng, m, "uniform")).astype(np.int32)
    assert np.array_equal(qo.operators.gatherImpl(r, loc), orc.gather_records(r, loc))
    perm = rng.permutation(n).astype(np.int32)  # distinct targets: the scatter's result is defined
    assert np.array_equal(qo.operators.scatterImpl(r, perm), orc.scatter_records(r, perm))
    proj = np.stack([loc.view(np.uint32), keys(rng, m, "uniform").view(np.uint32)], 1)
    assert np.array_equal(qo.operators.projection(r, proj), orc.project_records(r, proj))


@pytest.mark.parametrize("seed", range(12))
def test_random_planner_covers(qo, orc, seed):
    """A random filter -> group-by -> aggregate (or reduce, or selection) query
    through a random implementation cover, against the oracle's chain."""
    from query_on_opencl_fpga_amd import planner as P

    rng = np.random.default_rng(13_000 + seed)
    n = size(rng)
    G = int(rng.choice([1, 7, 1000, 5000]))
    r = np.stack([keys(rng, n, "uniform").view(np.uint32), rng.integers(0, G, n).astype(np.uint32)], 1)
    filt = bool(rng.random() < 0.7)
    lo, hi = sorted(int(x) for x in rng.integers(-10, G + 10, 2))
    if seed % 4 == 3:
        lo, hi = hi + 1, lo  # an empty range
    group = bool(rng.random() < 0.6)
    op = int(rng.integers(0, 4)) if (group or rng.random() < 0.7) else -1
    if not filt and not group and op < 0:
        filt = True
    q = P.Query(filter=filt, lo=lo, hi=hi, group=group, op=op, selectivity=0.5, groups=G)
    mask = int(rng.integers(0, 1 << 5))
    try:
        p = P.plan(q, n, mask)
    except qo._lib.QoError:
        p = P.plan(q, n)  # the mask allows no cover of this chain
    got = P.execute(p, q, r, group_capacity=G + 1)
    sel = orc.filter_records(r, lo, hi) if filt else r
    if group:
        srt = orc.sort_records(sel)
        sp = orc.group_startpos(srt)
        ek = srt[sp, 1] if len(sp) else np.zeros(0, np.uint32)
        assert np.array_equal(got[0], ek), p.describe()
        assert np.array_equal(got[1], orc.parallel_aggregate(srt, sp, op)), p.describe()
    elif op >= 0:
        assert got == orc.reduce_records(sel, op), p.describe()
    else:
        assert np.array_equal(got, sel), p.describe()


@pytest.mark.parametrize("seed", range(12))
def test_random_compat_chain(qo, orc, seed):
    """The reference-named host entry points on random Records: range selection,
    radix sort, group-by (sorted Records + start positions) and the four
    after-group-by aggregates of .x per group."""
    rng = np.random.default_rng(14_000 + seed)
    n = size(rng)
    r = np.stack([keys(rng, n, "uniform").view(np.uint32), keys(rng, n, KINDS[seed % len(KINDS)]).view(np.uint32)],
                 1)
    lo, hi = bounds(rng)
    assert np.array_equal(qo.operators.CL_RangeSelection(r, lo, hi), orc.filter_records(r, lo, hi))
    srt = orc.sort_records(r)
    assert np.array_equal(qo.operators.CL_RadixSort(r), srt)
    got_srt, sp = qo.operators.CL_GroupBy(r)
    assert np.array_equal(got_srt, srt) and np.array_equal(sp, orc.group_startpos(srt))
    for name, code in (("sum", 0), ("max", 1), ("min", 2), ("avg", 3)):
        got = getattr(qo.operators, f"CL_agg_{name}_afterGroupBy")(srt, sp)
        assert np.array_equal(got, orc.parallel_aggregate(srt, sp, code)), name


@pytest.mark.parametrize("seed", range(4))
def test_random_large_join_and_groupby(qo, orc, seed):
    """Sizes that take the partitioned paths: the join's sliced key-range or
    partitioned hash build and its L2-slice probe (|R| >= 2^20), the group-by's
    one-pass partition (groups beyond one LDS table), a 2-pass split."""
    rng = np.random.default_rng(15_000 + seed)
    nr = (1 << 21) + int(rng.integers(0, 1 << 20))
    ns = (1 << 22) + int(rng.integers(0, 1 << 21))
    if seed % 2 == 0:
        rk = (int(rng.choice([0, -nr, INT_MAX - nr + 1])) + rng.permutation(nr)).astype(np.int64).astype(np.int32)
    else:
        rk = rng.choice(np.arange(INT_MIN, INT_MAX, 997, dtype=np.int64), nr, replace=False).astype(np.int32)
    rp = keys(rng, nr, "uniform")
    sk = np.where(rng.random(ns) < 0.9, rk[rng.integers(0, nr, ns)], keys(rng, ns, "uniform")).astype(np.int32)
    sp = keys(rng, ns, "uniform")
    sf = rng.integers(0, 100, ns).astype(np.int32)
    got = qo.operators.hash_join(rk, rp, sk, sp, sf, INT_MIN, 59)
    exp = orc.hash_join(rk, rp, sk, sp, sf, INT_MIN, 59)
    assert got[3] == len(exp[0])
    assert np.array_equal(canon(np.stack(got[:3], 1)), canon(np.stack(exp, 1)))
    n = ns
    G = int(rng.choice([1 << 16, 1 << 20]))
    k = (rng.integers(0, G, n) * 2654435761 % (1 << 32)).astype(np.uint32).view(np.int32)
    v = keys(rng, n, "uniform")
    gg = qo.operators.groupby_agg(k, v, G)
    for a, b in zip(gg, orc.groupby_agg(k, v)):
        assert np.array_equal(a, b)
    r = np.stack([np.arange(n, dtype=np.uint32), rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)], 1)
    parts = int(rng.choice([300, 4097, 65536]))
    out, bound = qo.operators.splitImpl(r, parts)
    eo, eb = orc.split_records(r, parts)
    assert np.array_equal(out, eo) and np.array_equal(bound, eb)
