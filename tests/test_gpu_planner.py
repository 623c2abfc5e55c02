"""GPU: plans of the cost-model planner executed on the device against the
oracle's chain filter -> sort -> group start positions -> parallel aggregate
(the reference's group-by pipeline, host/host/testAggAfterGB.cpp:313-427 and
Validate.cpp:72-164) or filter -> reduce (Validate.cpp:372-468), for every
implementation cover the planner can choose."""
import ctypes as C

import numpy as np
import pytest

from query_on_opencl_fpga_amd import _lib
from query_on_opencl_fpga_amd import planner as P

pytestmark = pytest.mark.gpu

# Implementation sets that force each cover of each chain.
COVERS_GROUP_FILTERED = [(1 << P.IMPL_FILTER) | (1 << P.IMPL_SORT_GROUP_AGG),
                         (1 << P.IMPL_FILTER) | (1 << P.IMPL_HASH_GROUP_AGG),
                         1 << P.IMPL_FILTER_HASH_GROUP_AGG, 0]
COVERS_GROUP = [1 << P.IMPL_SORT_GROUP_AGG, 1 << P.IMPL_HASH_GROUP_AGG, 0]


def records(orc, n, seed, kmax, high_keys=False):
    """.x uniform 32-bit (negative as int32 too), .y uniform in [0, kmax); with
    high_keys some .y at and above 2^31 (negative as int32)."""
    x = orc.gen_uniform_u32(n, seed, 0, 0, 0)
    y = orc.gen_uniform_u32(n, seed, 1, 0, kmax)
    if high_keys:
        y = np.where(np.arange(n) % 7 == 3, y + np.uint32(0xFFFFFF00), y).astype(np.uint32)
    return np.stack([x, y], 1)


def expect_group(orc, r, q):
    sel = orc.filter_records(r, q.lo, q.hi) if q.filter else r
    srt = orc.sort_records(sel)
    sp = orc.group_startpos(srt)
    return srt[sp, 1] if len(sp) else np.zeros(0, np.uint32), orc.parallel_aggregate(srt, sp, q.op)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("mask", COVERS_GROUP_FILTERED)
def test_filtered_group_aggregate(qo, orc, op, mask):
    n = 300_007
    r = records(orc, n, 40 + op, 200)
    q = P.Query(filter=True, lo=10, hi=120, group=True, op=op, selectivity=0.55, groups=111)
    p = P.plan(q, n, mask)
    keys, aggs = P.execute(p, q, r, group_capacity=4096)
    ek, ea = expect_group(orc, r, q)
    assert np.array_equal(keys, ek), p.describe()
    assert np.array_equal(aggs, ea), p.describe()


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("mask", COVERS_GROUP)
@pytest.mark.parametrize("kmax", [1000, 1 << 20])
def test_group_aggregate_with_high_keys(qo, orc, op, mask, kmax):
    """Keys with the sign bit set order last (the reference sorts .y unsigned)."""
    n = 1 << 20
    r = records(orc, n, 50 + op, kmax, high_keys=True)
    q = P.Query(group=True, op=op, groups=min(kmax, n) + 256)
    p = P.plan(q, n, mask)
    keys, aggs = P.execute(p, q, r, group_capacity=(1 << 20) + 512)
    ek, ea = expect_group(orc, r, q)
    assert np.array_equal(keys, ek), p.describe()
    assert np.array_equal(aggs, ea), p.describe()


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("filtered", [False, True])
def test_reduce_and_selection(qo, orc, op, filtered):
    n = 1_000_003
    r = records(orc, n, 60 + op, 1 << 30)
    q = P.Query(filter=filtered, lo=0, hi=1 << 28, op=op, selectivity=0.25)
    got = P.execute(P.plan(q, n), q, r)
    sel = orc.filter_records(r, q.lo, q.hi) if filtered else r
    assert got == orc.reduce_records(sel, op)
    qs = P.Query(filter=True, lo=5, hi=1 << 20, op=-1, selectivity=0.001)
    assert np.array_equal(P.execute(P.plan(qs, n), qs, r), orc.filter_records(r, 5, 1 << 20))


def test_planner_choice_runs_at_scale(qo, orc):
    """The planner's own choice on 2^24 rows (1024 groups, half selected)."""
    n = 1 << 24
    r = records(orc, n, 70, 2048)
    q = P.Query(filter=True, lo=0, hi=1023, group=True, op=0, selectivity=0.5, groups=1024)
    p = P.plan(q, n)
    keys, aggs = P.execute(p, q, r, group_capacity=2048)
    ek, ea = expect_group(orc, r, q)
    assert np.array_equal(keys, ek) and np.array_equal(aggs, ea), p.describe()


def test_capacity_and_empty(qo, orc):
    r = records(orc, 10_000, 80, 5000)
    q = P.Query(group=True, op=0, groups=5000)
    for mask in COVERS_GROUP[:2]:
        with pytest.raises(_lib.QoError):
            P.execute(P.plan(q, 10_000, mask), q, r, group_capacity=100)
    empty = np.zeros((0, 2), np.uint32)
    for mask in COVERS_GROUP_FILTERED:
        qf = P.Query(filter=True, lo=0, hi=10, group=True, op=1, selectivity=0.5, groups=10)
        keys, aggs = P.execute(P.plan(qf, 0, mask), qf, empty, group_capacity=16)
        assert keys.size == 0 and aggs.size == 0


def test_calibrate_measures_every_primitive(qo):
    saved = P.cost_table()
    try:
        table = P.calibrate(20, 3)
        assert len(table) == len(P.COST_NAMES)
        assert all(f >= 0 and g > 0 for f, g in table)
        assert table != saved  # measured on this device, not the defaults
        q = P.Query(filter=True, lo=0, hi=49, group=True, op=0, selectivity=0.5, groups=50)
        assert P.plan(q, 1 << 24).est_ms > 0
    finally:
        for i, (f, g) in enumerate(saved):
            P.set_cost(i, f, g)
