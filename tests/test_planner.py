"""CPU: the cost-model planner's dynamic programme (qo_plan_query is host-only
code in libqo).  The plan it returns for the full implementation set must cost
the minimum over every cover that a restricted implementation set allows (each
restricted plan is a cover the library itself costs), cover the query's steps
in order, and follow the primitive cost table when that changes."""
import ctypes as C
import itertools

import pytest

from query_on_opencl_fpga_amd import _lib
from query_on_opencl_fpga_amd import planner as P

QUERIES = [
    P.Query(filter=True, lo=0, hi=49, group=True, op=0, selectivity=0.5, groups=50),
    P.Query(filter=True, lo=0, hi=49, group=True, op=3, selectivity=0.01, groups=1 << 20),
    P.Query(filter=True, lo=0, hi=9, group=True, op=1, selectivity=0.9, groups=1 << 20),
    P.Query(group=True, op=2, groups=1 << 10),
    P.Query(group=True, op=2, groups=1 << 22),
    P.Query(filter=True, lo=0, hi=5, op=1, selectivity=0.2),
    P.Query(op=0),
    P.Query(filter=True, lo=3, hi=3, op=-1, selectivity=0.001),
]
SIZES = [1 << 12, 1 << 20, 1 << 26, 1 << 28]


def chain_len(q):
    return int(q.filter) + int(q.group or q.op >= 0)


@pytest.fixture(autouse=True)
def keep_costs():
    saved = P.cost_table()
    yield
    for i, (f, g) in enumerate(saved):
        P.set_cost(i, f, g)


@pytest.mark.parametrize("qi", range(len(QUERIES)))
@pytest.mark.parametrize("n", SIZES)
def test_plan_is_the_cheapest_cover(qi, n):
    q = QUERIES[qi]
    best = P.plan(q, n)
    costs = []
    for k in range(1, 6):
        for impls in itertools.combinations(range(5), k):
            mask = sum(1 << i for i in impls)
            try:
                p = P.plan(q, n, mask)
            except _lib.QoError:
                continue  # no cover with these implementations
            assert set(p.impls) <= set(impls)
            costs.append(p.est_ms)
    assert costs and best.est_ms == pytest.approx(min(costs), rel=1e-12)
    # The segments cover steps 0..L-1 in order, and the estimate is their sum.
    assert best.steps[0][0] == 0 and best.steps[-1][1] == chain_len(q) - 1
    for (a0, b0), (a1, b1) in zip(best.steps, best.steps[1:]):
        assert a1 == b0 + 1
    assert best.est_ms == pytest.approx(sum(best.seg_ms), rel=1e-12)


def test_plan_follows_the_cost_table():
    q = P.Query(filter=True, lo=0, hi=49, group=True, op=0, selectivity=0.5, groups=1 << 20)
    n = 1 << 26
    # Hash group-by (plain and filtered) made very slow: the sort-based pipeline wins.
    for prim in (5, 6, 8, 9):
        P.set_cost(prim, 10.0, 1.0)
    p = P.plan(q, n)
    assert p.impls == [P.IMPL_FILTER, P.IMPL_SORT_GROUP_AGG]
    # Radix passes made very slow instead: a hash implementation wins.
    for prim in (5, 6, 8, 9):
        P.set_cost(prim, 10.0, 4000.0)
    P.set_cost(1, 10.0, 1.0)
    assert P.IMPL_SORT_GROUP_AGG not in P.plan(q, n).impls
    # Filter compaction made very slow: the fused filter + hash group-by wins.
    P.set_cost(0, 1e6, 1.0)
    assert P.plan(q, n).impls == [P.IMPL_FILTER_HASH_GROUP_AGG]


def test_plan_reports_launch_parameters_from_the_tune_table():
    q = P.Query(group=True, op=0, groups=1024)
    p = P.plan(q, 1 << 24, 1 << P.IMPL_HASH_GROUP_AGG)
    op, block, items, param = p.launch[0]
    assert op == 3  # the direct group-by's tune operator
    b, it, pa = C.c_int(), C.c_int(), C.c_int()
    _lib.call("qo_tune_get", 3, 24, C.byref(b), C.byref(it), C.byref(pa))
    assert (block, items, param) == (b.value, it.value, pa.value)


def test_plan_errors():
    with pytest.raises(_lib.QoError):
        P.plan(P.Query(group=True, op=-1), 100)           # a group needs an aggregate
    with pytest.raises(_lib.QoError):
        P.plan(P.Query(op=7), 100)                         # no such operator
    with pytest.raises(_lib.QoError):
        P.plan(P.Query(filter=True, selectivity=1.5, op=0), 100)
    with pytest.raises(_lib.QoError):
        P.plan(P.Query(), 100)                             # nothing to do
    with pytest.raises(_lib.QoError):
        P.plan(P.Query(group=True, op=0), 100, 1 << P.IMPL_REDUCE)  # no cover
    with pytest.raises(_lib.QoError):
        P.set_cost(99, 1.0, 1.0)
    with pytest.raises(_lib.QoError):
        P.set_cost(0, 1.0, 0.0)
