"""GPU parity at the benchmark configurations' scale.

* Group-by, (G, z) in {2^10, 2^20} x {0, 0.5, 1}: exact against the oracle at
  2M tuples, and at the full 2^28 tuples through size-independent properties
  (sum of counts = N, sum of sums = sum of values, keys strictly ascending, at
  most G groups) plus the oracle's aggregates of 1024 sampled keys recomputed
  from their tuples.
* Batched hj_filter_sort: 512 small queries (|R| log-uniform in [4K, 64K]),
  every query exact against the oracle; the full 512-query configuration
  (|R| log-uniform in [4K, 4M]) with every query's match count equal to its
  filtered S count, its output sorted, and its order-sensitive digest equal to
  the plain path's.
* Filtered hash join at R = 2^24, S = 2^28: match count = filtered S count, the
  multiset of output rows equal to the one the inputs define (order-independent
  hash over all rows), and 1M sampled output rows whose r_payload is the
  payload of their key in R.
"""
import ctypes as C

import numpy as np
import pytest

from query_on_opencl_fpga_amd import data
from query_on_opencl_fpga_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1
CASES = [(1 << 10, 0.0), (1 << 10, 0.5), (1 << 10, 1.0), (1 << 20, 0.0), (1 << 20, 0.5), (1 << 20, 1.0)]


@pytest.mark.parametrize("G,z", CASES)
def test_groupby_config_cases_match_oracle(qo, orc, G, z):
    n = 2_000_000
    t = orc.zipf_table(G, z)
    keys = orc.gen_zipf_keys(n, t, 13, 1, 29).astype(np.int32)
    vals = orc.gen_uniform_u32(n, 13, 2, 0, 1 << 16).astype(np.int32)
    got = qo.operators.groupby_agg(keys, vals, G)
    exp = orc.groupby_agg(keys, vals)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def _mix(x):
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


@pytest.mark.parametrize("G,z", CASES)
def test_groupby_full_size_properties(qo, orc, G, z):
    n = 1 << 28
    print(f"[scale] group-by n=2^28 G={G} z={z}", flush=True)
    t = data.GroupbyTables(n, G, z, 0x51A7E5EED0000003)
    outs = [DeviceBuffer(G * s) for s in (4, 8, 8, 4, 4)]
    num = DeviceBuffer(8)
    tot = DeviceBuffer(24)
    try:
        qo._lib.call("qo_groupby_agg", t.key.ptr, t.val.ptr, n, G, *[b.ptr for b in outs], num.ptr, None)
        g = int(num.to_numpy(np.uint64, 1)[0])
        assert 1 <= g <= G
        key = outs[0].to_numpy(np.int32, g)
        cnt, sm = outs[1].to_numpy(np.int64, g), outs[2].to_numpy(np.int64, g)
        mn, mx = outs[3].to_numpy(np.int32, g), outs[4].to_numpy(np.int32, g)
        assert np.all(np.diff(key) > 0)
        assert int(cnt.sum()) == n and cnt.min() >= 1
        qo._lib.call("qo_filter_agg", t.val.ptr, t.val.ptr, n, INT_MIN, INT_MAX, tot.ptr, None)
        assert int(sm.sum()) == int(tot.to_numpy(np.int64, 3)[1])
        # The oracle's aggregates for 1024 sampled keys, from their tuples.
        hk, hv = t.key.to_numpy(np.int32, n), t.val.to_numpy(np.int32, n)
        rng = np.random.default_rng(G + int(10 * z))
        pick = rng.choice(g, size=min(1024, g), replace=False)
        sel = np.isin(hk, key[pick])
        exp = orc.groupby_agg(hk[sel], hv[sel])
        idx = np.searchsorted(key, exp[0])
        assert np.array_equal(key[idx], exp[0])
        for col, e in zip((cnt, sm, mn, mx), exp[1:]):
            assert np.array_equal(col[idx], e)
    finally:
        for b in outs + [num, tot]:
            b.free()
        t.free()


def _batch_queries(qo, nq, scale, seed):
    """Config-4 queries (the benchmark's generator): |R| = 2^(12 + 10*scale*u),
    unique 31-bit keys, |S| = |R| * U[4, 16], fk uniform over R's keys, filter
    U[0, 1000) < 1000 * U[0.05, 0.5]."""
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    qs = []
    for i in range(nq):
        nr = int(2 ** (12 + 10 * rng.random() * scale))
        ns = int(nr * (4 + 12 * rng.random()))
        fhi = int(1000 * (0.05 + 0.45 * rng.random())) - 1
        s = seed + 1000 + i
        q = {"nr": nr, "ns": ns, "flo": INT_MIN, "fhi": fhi, "rkey": data.gen_bij(nr, 31, s),
             "rpay": data.gen_uniform(nr, s, 1, 0, 0), "skey": data.gen_fk(ns, s, 2, nr, 31, s),
             "spay": data.gen_uniform(ns, s, 3, 0, 0), "sfilt": data.gen_uniform(ns, s, 4, 0, 1000)}
        q["cap"] = int(ns * (fhi + 1) / 1000 * 1.1) + 4096
        for c in ("okey", "orpay", "ospay"):
            q[c] = DeviceBuffer(q["cap"] * 4)
        qs.append(q)
    qo._lib.call("qo_device_sync")
    return qs


def _free_queries(qs):
    for q in qs:
        for v in q.values():
            if hasattr(v, "free"):
                v.free()


def test_batch_512_small_queries_match_oracle(qo, orc):
    print("[scale] 512-query batch, |R| in [4K, 64K], vs oracle", flush=True)
    qs = _batch_queries(qo, 512, 0.4, 0x51A7E5EED0000004)
    b = qo.operators.JoinBatch(qs)
    try:
        b.launch()
        counts = b.counts()
        for i, (q, m) in enumerate(zip(qs, counts)):
            m = int(m)
            rk, rp = q["rkey"].to_numpy(np.int32, q["nr"]), q["rpay"].to_numpy(np.int32, q["nr"])
            sk, sp, sf = (q[c].to_numpy(np.int32, q["ns"]) for c in ("skey", "spay", "sfilt"))
            exp = orc.sort_join_rows(*orc.hash_join(rk, rp, sk, sp, sf, q["flo"], q["fhi"]))
            assert m == len(exp[0]), i
            for c, e in zip(("okey", "orpay", "ospay"), exp):
                assert np.array_equal(q[c].to_numpy(np.int32, m), e), (i, c)
    finally:
        b.destroy()
        _free_queries(qs)


def test_batch_full_config_every_query(qo):
    print("[scale] 512-query batch, |R| in [4K, 4M]: counts, order, digests vs plain", flush=True)
    lib = qo._lib.load()
    qs = _batch_queries(qo, 512, 1.0, 0x51A7E5EED0000004)
    dig = DeviceBuffer(24)
    fc = DeviceBuffer(24)

    def digest(q, m):
        qo._lib.call("qo_join_rows_digest", q["okey"].ptr, q["orpay"].ptr, q["ospay"].ptr, int(m), None,
                     dig.ptr, None)
        return tuple(int(x) for x in dig.to_numpy(np.uint64, 3))

    b = qo.operators.JoinBatch(qs)
    try:
        b.launch()
        counts = [int(c) for c in b.counts()]
        batched = [digest(q, m) for q, m in zip(qs, counts)]
    finally:
        b.destroy()
    try:
        table = DeviceBuffer(lib.qo_join_table_bytes(max(q["nr"] for q in qs)))
        m = C.c_uint64()
        for i, q in enumerate(qs):
            qo._lib.call("qo_filter_agg", q["sfilt"].ptr, q["sfilt"].ptr, q["ns"], q["flo"], q["fhi"], fc.ptr, None)
            assert counts[i] == int(fc.to_numpy(np.int64, 3)[0]), i  # every fk is an R key
            assert batched[i][2] == 0, i  # sorted by (key, s_payload)
            qo._lib.call("qo_hj_filter_sort", table.ptr, lib.qo_join_table_bytes(q["nr"]), q["rkey"].ptr,
                         q["rpay"].ptr, q["nr"], q["skey"].ptr, q["spay"].ptr, q["sfilt"].ptr, q["ns"], q["flo"],
                         q["fhi"], q["okey"].ptr, q["orpay"].ptr, q["ospay"].ptr, q["cap"], C.byref(m), None)
            assert m.value == counts[i] and digest(q, m.value) == batched[i], i
        table.free()
    finally:
        dig.free()
        fc.free()
        _free_queries(qs)


def test_join_full_size_config(qo):
    print("[scale] join R=2^24 S=2^28: count, multiset hash, sampled payloads", flush=True)
    lib = qo._lib.load()
    nr, ns = 1 << 24, 1 << 28
    t = data.JoinTables(nr, ns, 0x51A7E5EED0000002)
    tb = lib.qo_join_table_bytes(nr)
    table = DeviceBuffer(tb)
    outs = [DeviceBuffer(ns * 4) for _ in range(3)]
    cnt = DeviceBuffer(8)
    fc = DeviceBuffer(24)
    try:
        qo._lib.call("qo_hash_build", table.ptr, tb, t.rkey.ptr, t.rpay.ptr, nr, None)
        qo._lib.call("qo_hash_probe", table.ptr, tb, t.skey.ptr, t.spay.ptr, t.sfilt.ptr, ns, INT_MIN, 49,
                     outs[0].ptr, outs[1].ptr, outs[2].ptr, ns, cnt.ptr, None)
        m = int(cnt.to_numpy(np.uint64, 1)[0])
        qo._lib.call("qo_filter_agg", t.sfilt.ptr, t.sfilt.ptr, ns, INT_MIN, 49, fc.ptr, None)
        assert m == int(fc.to_numpy(np.int64, 3)[0])
        ok, orp, osp = (o.to_numpy(np.int32, m) for o in outs)
        rk, rp = t.rkey.to_numpy(np.int32, nr), t.rpay.to_numpy(np.int32, nr)
        pay_of_key = np.empty(nr, np.int32)
        pay_of_key[rk] = rp  # R's keys are a permutation of [0, nr)
        # 1M sampled output rows: r_payload is the payload of the row's key in R.
        pick = np.random.default_rng(5).choice(m, size=1 << 20, replace=False)
        assert np.array_equal(orp[pick], pay_of_key[ok[pick]])
        # All rows: order-independent hash of (key, r_payload, s_payload) equals
        # the hash of the rows the filtered S defines.
        sk, sp, sf = (c.to_numpy(np.int32, ns) for c in (t.skey, t.spay, t.sfilt))
        sel = sf <= 49
        ek, es = sk[sel], sp[sel]
        del sk, sp, sf, sel

        def rows_hash(k, r, s):
            a = (k.astype(np.uint64) << np.uint64(32)) | r.astype(np.uint32).astype(np.uint64)
            with np.errstate(over="ignore"):
                return int(np.sum(_mix(_mix(a) ^ s.astype(np.uint32).astype(np.uint64)), dtype=np.uint64))

        assert rows_hash(ok, orp, osp) == rows_hash(ek, pay_of_key[ek], es)
    finally:
        for b in outs + [table, cnt, fc]:
            b.free()
        t.free()
