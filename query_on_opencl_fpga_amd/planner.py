"""Cost-model planner for filter -> group-by -> aggregate queries over Records.

The analogue of the reference's FPGA planner (host/dynamic_programming.cpp:
874-1023): a query's chain of steps (filter, group + aggregate, or reduce) is
covered by the cheapest sequence of libqo implementations, found by a dynamic
programme over the chain's prefixes in libqo (``qo_plan_query``), with each
implementation costed from a table of primitive costs (fixed microseconds +
bytes / bandwidth) that ``calibrate`` measures on the device.  ``execute`` runs
a plan on the GPU.

    q = Query(filter=True, lo=0, hi=49, group=True, op=REDUCE_SUM, selectivity=0.5, groups=50)
    p = plan(q, n)                      # host only
    keys, aggs = execute(p, q, records, group_capacity=1024)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceBuffer

IMPL_FILTER, IMPL_SORT_GROUP_AGG, IMPL_HASH_GROUP_AGG, IMPL_FILTER_HASH_GROUP_AGG, IMPL_REDUCE = range(5)
IMPL_NAMES = ["filter (one-pass compaction)", "sort + startPos + aggregate", "hash group-by",
              "filtered hash group-by (fused)", "reduce"]
COST_NAMES = ["filter", "sort pass", "group start positions", "per-group aggregate", "Records -> columns",
              "hash group-by (<= 1024 groups)", "hash group-by (partitioned)", "reduce",
              "filtered hash group-by, half selected (<= 1024 groups)",
              "filtered hash group-by, half selected (partitioned)"]


@dataclass
class Query:
    filter: bool = False
    lo: int = 0
    hi: int = 0
    group: bool = False
    op: int = -1            # REDUCE_SUM / MAX / MIN / AVERAGE = 0..3; -1: selection only
    selectivity: float = 1.0
    groups: float = 1.0

    def c(self) -> _lib.QoQuery:
        return _lib.QoQuery(int(self.filter), int(self.lo), int(self.hi), int(self.group), int(self.op),
                            float(self.selectivity), float(self.groups))


@dataclass
class Plan:
    impls: list
    steps: list             # (first, last) chain positions per segment
    seg_ms: list
    est_ms: float
    launch: list            # (tune op, block, items, param) per segment
    raw: _lib.QoPlan

    def describe(self) -> str:
        segs = ", ".join(f"{IMPL_NAMES[i]} [{a}..{b}] {t:.3f} ms" for i, (a, b), t in
                         zip(self.impls, self.steps, self.seg_ms))
        return f"{segs}; total {self.est_ms:.3f} ms"


def plan(q: Query, n: int, allowed: int = 0) -> Plan:
    """Cheapest plan for n rows using the implementations in the `allowed` bit mask (0 = all)."""
    p = _lib.QoPlan()
    qc = q.c()
    _lib.call("qo_plan_query", C.byref(qc), int(n), int(allowed), C.byref(p))
    k = p.nseg
    return Plan([p.impl[i] for i in range(k)], [(p.first_step[i], p.last_step[i]) for i in range(k)],
                [p.seg_ms[i] for i in range(k)], p.est_ms,
                [(p.tune_op[i], p.block[i], p.items[i], p.param[i]) for i in range(k)], p)


def cost_table() -> list:
    out = []
    for i in range(len(COST_NAMES)):
        f, g = C.c_double(), C.c_double()
        _lib.call("qo_plan_cost_get", i, C.byref(f), C.byref(g))
        out.append((f.value, g.value))
    return out


def set_cost(prim: int, fixed_us: float, gbps: float) -> None:
    _lib.call("qo_plan_cost_set", int(prim), float(fixed_us), float(gbps))


def calibrate(size_class: int = 24, reps: int = 5) -> list:
    """Measure every primitive on the device at 2^size_class rows; returns the new table."""
    _lib.call("qo_plan_calibrate", int(size_class), int(reps))
    return cost_table()


def execute(p: Plan, q: Query, records, group_capacity: int = 0):
    """Run plan p on host Records (numpy (n, 2) uint32).  Returns the selected
    Records (selection), (keys, aggregates) (grouped) or the reduce result."""
    r = np.ascontiguousarray(records, dtype=np.uint32).reshape(-1, 2)
    n = r.shape[0]
    cap = max(int(group_capacity), 1)
    qc = q.c()
    with DeviceBuffer.from_numpy(r) as d_in, DeviceBuffer(max(n, 1) * 8) as d_sel, \
            DeviceBuffer(cap * 4) as d_keys, DeviceBuffer(cap * 4) as d_aggs, DeviceBuffer(8) as d_cnt:
        _lib.call("qo_plan_execute", C.byref(p.raw), C.byref(qc), d_in.ptr, n, int(group_capacity),
                  d_sel.ptr, d_keys.ptr, d_aggs.ptr, d_cnt.ptr, None)
        m = int(d_cnt.to_numpy(np.uint64, 1)[0])
        if q.group:
            return d_keys.to_numpy(np.uint32, m), d_aggs.to_numpy(np.int32, m)
        if q.op >= 0:
            return int(d_aggs.to_numpy(np.uint32, 1)[0])
        return d_sel.to_numpy(np.uint32, 2 * m).reshape(m, 2)
