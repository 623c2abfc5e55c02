#!/usr/bin/env python3
"""Benchmark of the relational operator path on one MI355X.

Headline workload (one step = one pass of the hot path over one batch):
filtered hash join, R = 2^24 tuples (key = seeded permutation of [0, 2^24),
payload uniform int32), S = 2^28 tuples (fk uniform over R keys, payload
uniform int32, filter column uniform in [0, 100) with predicate < 50); a step
builds the table on R and probes it with the filtered S, materialising
(key, r_payload, s_payload).  Inputs are generated in HBM before timing.

The default run also measures the other configurations and reports each
under "configs", every one with its roofline and CPU baseline: fused filter +
aggregation at 2^28 tuples, hash group-by at 2^28 tuples for G in {2^10, 2^20}
x Zipf z in {0, 0.5, 1}, and batched hj_filter_sort (512 queries in one graph
launch against the plain per-query path) at the configured sizes and for
small queries.  --headline-only skips them.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--config join|filter_agg|groupby|batched|served|copy|ops|plan] [--headline-only]

Prints ONE JSON line.  `value` is S tuples per second of the whole join step
(build + probe); the roofline block covers the probe stage, timed with HIP
events on the stream it runs on.  CPU baselines run the multi-threaded CPU
restatement of the reference algorithms (oracle/qo_cpu_par.c) on the host's
cores on the same inputs (copied back from HBM), median of 5 runs after a
warm-up.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "probe/scan tuples/s + % HBM peak; batched-vs-plain queries/s; p99 ms"
SEED_BASE = 0x51A7E5EED0000000
INT_MIN = -(1 << 31)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profile_traffic(name: str, match: dict):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE)
    for this workload, from profiles/<name>_traffic.json when it was measured on
    the same configuration; None otherwise."""
    path = ROOT / "profiles" / f"{name}_traffic.json"
    try:
        rec = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    if any(rec.get(k) != v for k, v in match.items()):
        return None
    return rec.get("traffic_bytes")


# ---------------------------------------------------------------------------
# CPU baseline: the multi-threaded restatement of the reference algorithms
# (oracle/qo_cpu_par.c) on this host's cores, on the GPU run's own inputs.
# ---------------------------------------------------------------------------
CPU_RUNS = 5


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the cores this process may use, at
    most 16 (one GPU's share of the host on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    return orc, orc.load()


def _median_time(fn) -> tuple[float, list]:
    """Median wall time of fn() over CPU_RUNS runs after one warm-up run."""
    fn()
    ts = []
    for _ in range(CPU_RUNS):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def _baseline(value, unit, sample, times, extra=None) -> dict:
    d = {"value": value, "unit": unit, "cores": cpu_threads(), "kind": "port", "sample": sample,
         "cpu_model": cpu_model(), "runs": CPU_RUNS, "median_s": float(np.median(times)),
         "timer": "time.perf_counter around each call (CLOCK_MONOTONIC), data generation and copies excluded"}
    if extra:
        d.update(extra)
    return d


def cpu_baseline_join(rk, rp, sk, sp, sf, flo, fhi) -> dict:
    """Filtered hash join (bucket-chained build on R, probe of filtered S,
    materialised rows) on the full inputs."""
    orc, lib = _oracle()
    nt = cpu_threads()
    nr, ns = rk.shape[0], sk.shape[0]
    m = lib.pr_hash_join(rk.ctypes.data, rp.ctypes.data, nr, sk.ctypes.data, sp.ctypes.data, sf.ctypes.data, ns,
                         flo, fhi, None, None, None, 0, nt)
    outs = [np.empty(max(m, 1), np.int32) for _ in range(3)]

    def run():
        lib.pr_hash_join(rk.ctypes.data, rp.ctypes.data, nr, sk.ctypes.data, sp.ctypes.data, sf.ctypes.data, ns,
                         flo, fhi, outs[0].ctypes.data, outs[1].ctypes.data, outs[2].ctypes.data, m, nt)

    t, ts = _median_time(run)
    return _baseline(ns / t, "S tuples/s", f"full workload: build on |R|={nr}, filtered probe of |S|={ns}, "
                     f"{m} rows materialised; pr_hash_join on {nt} threads", ts, {"matches": int(m)})


def cpu_baseline_filter_agg(k, v, lo, hi) -> dict:
    orc, lib = _oracle()
    nt = cpu_threads()
    out = np.zeros(3, np.int64)
    t, ts = _median_time(lambda: lib.pr_filter_agg(k.ctypes.data, v.ctypes.data, k.shape[0], lo, hi,
                                                   out.ctypes.data, nt))
    return _baseline(k.shape[0] / t, "tuples/s", f"full workload: {k.shape[0]} tuples; pr_filter_agg on {nt} "
                     "threads", ts)


def cpu_baseline_groupby(k, v, note) -> dict:
    """The reference's sort-based group-by (sort -> start positions -> per-group
    aggregates) on a bounded sample of the case's tuples."""
    orc, lib = _oracle()
    nt = cpu_threads()
    n = k.shape[0]
    outs = [np.empty(n, d) for d in (np.int32, np.int64, np.int64, np.int32, np.int32)]
    t, ts = _median_time(lambda: lib.pr_groupby_agg(k.ctypes.data, v.ctypes.data, n, *[o.ctypes.data for o in outs],
                                                    nt))
    return _baseline(n / t, "tuples/s", f"{note}; pr_groupby_agg (parallel LSD radix sort + group reduce) on "
                     f"{nt} threads", ts)


def cpu_baseline_batch(queries) -> dict:
    """The plain per-query path on the CPU: filtered join, then sort by
    (key, s_payload), one query after the other, for a sample of the queries
    (each query's inputs copied back from HBM)."""
    orc, lib = _oracle()
    nt = cpu_threads()
    bufs = []
    for q in queries:
        rk, rp = q["rkey"].to_numpy(np.int32, q["nr"]), q["rpay"].to_numpy(np.int32, q["nr"])
        sk, sp, sf = (q[c].to_numpy(np.int32, q["ns"]) for c in ("skey", "spay", "sfilt"))
        cap = q["cap"]
        bufs.append((rk, rp, sk, sp, sf, q["flo"], q["fhi"], [np.empty(cap, np.int32) for _ in range(3)], cap))

    def run():
        for rk, rp, sk, sp, sf, flo, fhi, o, cap in bufs:
            m = lib.pr_hash_join(rk.ctypes.data, rp.ctypes.data, rk.shape[0], sk.ctypes.data, sp.ctypes.data,
                                 sf.ctypes.data, sk.shape[0], flo, fhi, o[0].ctypes.data, o[1].ctypes.data,
                                 o[2].ctypes.data, cap, nt)
            lib.pr_sort_join_rows(o[0].ctypes.data, o[1].ctypes.data, o[2].ctypes.data, min(m, cap), nt)

    t, ts = _median_time(run)
    return _baseline(len(queries) / t, "queries/s", f"{len(queries)} queries of the batch (the first ones), each "
                     f"pr_hash_join + pr_sort_join_rows on {nt} threads, one after the other", ts)


# ---------------------------------------------------------------------------
def run_join(args, qo, lib) -> dict:
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Event, Stream

    nr, ns = args.nr, args.ns
    seed = SEED_BASE + 2
    log(f"[bench] join: generating R={nr} S={ns}")
    t = data.JoinTables(nr, ns, seed)
    tb = lib.qo_join_table_bytes(nr)
    if args.load:
        tb = 256 + 8 * (int(nr * 100 / args.load) + 1)
    table = DeviceBuffer(tb)
    cap = ns
    outs = [DeviceBuffer(cap * 4) for _ in range(3)]
    cnt = DeviceBuffer(8)
    st = Stream()
    evs = [[Event(), Event(), Event()] for _ in range(args.steps)]
    e_start, e_end = Event(), Event()

    def step(ev=None):
        if ev:
            ev[0].record(st)
        qo._lib.call("qo_hash_build", table.ptr, tb, t.rkey.ptr, t.rpay.ptr, nr, st.handle)
        if ev:
            ev[1].record(st)
        qo._lib.call("qo_hash_probe", table.ptr, tb, t.skey.ptr, t.spay.ptr, t.sfilt.ptr, ns, INT_MIN, 49,
                     outs[0].ptr, outs[1].ptr, outs[2].ptr, cap, cnt.ptr, st.handle)
        if ev:
            ev[2].record(st)

    log("[bench] join: warmup")
    for _ in range(args.warmup):
        step()
    st.sync()
    qo._lib.call("qo_device_sync")
    log("[bench] join: timed steps")
    e_start.record(st)
    for i in range(args.steps):
        step(evs[i])
    e_end.record(st)
    st.sync()
    total_ms = e_start.elapsed_ms(e_end)
    build_ms = [evs[i][0].elapsed_ms(evs[i][1]) for i in range(args.steps)]
    probe_ms = [evs[i][1].elapsed_ms(evs[i][2]) for i in range(args.steps)]
    m = int(cnt.to_numpy(np.uint64, 1)[0])
    # Size-independent check: every S fk is an R key, so matches = rows passing the filter.
    fcnt = DeviceBuffer(24)
    qo._lib.call("qo_filter_agg", t.sfilt.ptr, t.sfilt.ptr, ns, INT_MIN, 49, fcnt.ptr, None)
    expect = int(fcnt.to_numpy(np.int64, 3)[0])
    if m != expect:
        raise SystemExit(f"join produced {m} matches, expected {expect}")
    ms_step = total_ms / args.steps
    probe_avg = float(np.mean(probe_ms))
    build_avg = float(np.mean(build_ms))
    probe_bytes = 12 * ns + 12 * m
    achieved = probe_bytes / (probe_avg * 1e-3) / 1e9
    import ctypes as C
    mode, used, dups = C.c_int(), C.c_uint64(), C.c_int()
    qo._lib.call("qo_join_table_info", table.ptr, C.byref(mode), C.byref(used), C.byref(dups), st.handle)
    part, bits, splits = C.c_int(), C.c_int(), C.c_int()
    qo._lib.call("qo_join_strategy", tb, ns, mode.value, C.byref(part), C.byref(bits), C.byref(splits))
    table_kind = (f"key-range table ({used.value} slots, home slot = key - min key)" if mode.value == 1
                  else f"hash-mixed table ({used.value} slots)")
    # The key-range LDS route is taken on the device by tables of 2^21 slots and
    # more (the bench's keys have no (-1, -1) build tuple); smaller ones take L2.
    if part.value == 2 and mode.value == 1 and not dups.value and used.value >= (1 << 21):
        kernel = "probe stage (one qo_hash_probe call): k_range_lds_split + k_range_lds_probe"
        strategy = (f"{table_kind}; probe: one split by key into {used.value >> 15} buckets of 8 per-XCD "
                    "regions, lookups in LDS slices of 2^15 slots")
    elif part.value and splits.value == 1:
        kernel = "probe stage (one qo_hash_probe call): k_split_cols + k_l2_probe"
        strategy = f"{table_kind}; probe: one split into {1 << bits.value} buckets of 8 per-XCD regions, XCD-local L2 lookups"
    elif part.value:
        kernel = ("probe stage (one qo_hash_probe call): k_split_cols"
                  + (" + k_split_pairs" if splits.value == 2 else "") + " + k_slice_probe")
        strategy = f"{table_kind}; probe: {1 << bits.value} LDS slices, {splits.value} split pass(es)"
    else:
        kernel, strategy = "k_join_probe", f"{table_kind}; probe: direct"
    traffic, traffic_src = args.traffic, "--traffic on the command line"
    if traffic is None:
        traffic = profile_traffic("join_probe", {"R_tuples": nr, "S_tuples": ns, "strategy": strategy})
        traffic_src = ("stored profile profiles/join_probe_traffic.json (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE "
                       "of this configuration; not measured in this run)") if traffic is not None else None
    res = {
        "metric": METRIC,
        "value": ns / (ms_step * 1e-3),
        "unit": "S tuples/s",
        "ms_per_step": ms_step,
        "dtype": "int32",
        "config": {"workload": "filter + hash join (build R, probe filtered S, materialise key/r_payload/s_payload)",
                   "R_tuples": nr, "S_tuples": ns, "filter": "sfilt < 50 of U[0,100)", "matches": m,
                   "table_bytes": tb, "strategy": strategy},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "alg_bytes_per_launch": probe_bytes, "alg_bytes_formula": "12*|S| + 12*matches",
                     "avg_launch_ms": probe_avg,
                     "timed_with": "HIP events on the probe stream around each qo_hash_probe call"},
        "kernels": {"build_ms": build_avg, "probe_ms": probe_avg,
                    "build_R_tuples_per_s": nr / (build_avg * 1e-3),
                    "probe_S_tuples_per_s": ns / (probe_avg * 1e-3),
                    "build_GBps_alg": 16 * nr / (build_avg * 1e-3) / 1e9},
    }
    for b in outs + [table, cnt, fcnt]:
        b.free()
    if not args.no_cpu_baseline:
        log("[bench] join: cpu baseline (inputs copied back from HBM)")
        host = [t.rkey.to_numpy(np.int32, nr), t.rpay.to_numpy(np.int32, nr)] + \
               [c.to_numpy(np.int32, ns) for c in (t.skey, t.spay, t.sfilt)]
        t.free()
        res["cpu_baseline"] = cpu_baseline_join(*host, INT_MIN, 49)
        if res["cpu_baseline"]["matches"] != m:
            raise SystemExit(f"CPU baseline found {res['cpu_baseline']['matches']} matches, GPU {m}")
        del host
    else:
        t.free()
    return res


def run_filter_agg(args, qo, lib) -> dict:
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Event, Stream

    n = args.n_filter
    seed = SEED_BASE + 1
    log(f"[bench] filter_agg: generating n={n}")
    t = data.FilterAggTables(n, seed)
    out = DeviceBuffer(24)
    st = Stream()
    evs = [(Event(), Event()) for _ in range(args.steps)]
    for _ in range(args.warmup):
        qo._lib.call("qo_filter_agg", t.key.ptr, t.val.ptr, n, 0, (1 << 30) - 1, out.ptr, st.handle)
    st.sync()
    e0, e1 = Event(), Event()
    e0.record(st)
    for i in range(args.steps):
        evs[i][0].record(st)
        qo._lib.call("qo_filter_agg", t.key.ptr, t.val.ptr, n, 0, (1 << 30) - 1, out.ptr, st.handle)
        evs[i][1].record(st)
    e1.record(st)
    st.sync()
    total_ms = e0.elapsed_ms(e1)
    k_ms = float(np.mean([a.elapsed_ms(b) for a, b in evs]))
    r = out.to_numpy(np.int64, 3)
    achieved = 8 * n / (k_ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": n / (total_ms / args.steps * 1e-3), "unit": "tuples/s",
        "ms_per_step": total_ms / args.steps, "dtype": "int32",
        "config": {"workload": "fused filter + aggregation (COUNT, SUM int64) over key/value int32 columns",
                   "tuples": n, "predicate": "0 <= key <= 2^30-1", "count": int(r[0]), "sum": int(r[1])},
        "roofline": {"bound": "hbm", "kernel": "k_filter_agg", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": args.traffic if args.traffic is not None else profile_traffic("filter_agg", {"tuples": n}),
                     "alg_bytes_per_launch": 8 * n, "avg_launch_ms": k_ms},
    }
    out.free()
    if not args.no_cpu_baseline:
        log("[bench] filter_agg: cpu baseline")
        k, v = t.key.to_numpy(np.int32, n), t.val.to_numpy(np.int32, n)
        res["cpu_baseline"] = cpu_baseline_filter_agg(k, v, 0, (1 << 30) - 1)
        del k, v
    t.free()
    return res


def _time_op(st, steps, warmup, fn):
    """Average ms of fn() on stream st over `steps` calls (HIP events around each)."""
    from query_on_opencl_fpga_amd.device import Event
    for _ in range(warmup):
        fn()
    st.sync()
    evs = [(Event(), Event()) for _ in range(steps)]
    for a, b in evs:
        a.record(st)
        fn()
        b.record(st)
    st.sync()
    return float(np.mean([a.elapsed_ms(b) for a, b in evs]))


def run_ops(args, qo, lib) -> dict:
    """The remaining operators of the reference's path, each at a large size:
    filter compaction, exclusive scan, Record sort, the reference's sort-based
    group-by pipeline (sort -> startPos -> four aggregates), split, gather,
    scatter, projection, reduce, Records -> columns, and filter + join + sort
    (hj_filter_sort) on the join workload.  Each row: ms per call (HIP
    events on the call's stream) and algorithmic GB/s against the HBM peak."""
    import ctypes as C
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Stream

    st = Stream()
    rows = []

    def row(name, n, ms, alg, note):
        gbs = alg / (ms * 1e-3) / 1e9
        rows.append({"op": name, "n": n, "ms": ms, "alg_bytes": alg, "GBps_alg": gbs,
                     "frac_hbm": gbs / HBM_PEAK_GBS, "alg_bytes_formula": note})
        log(f"[bench] ops: {name} n={n}: {ms:.3f} ms, {gbs:.0f} GB/s alg")

    seed = SEED_BASE + 5
    # Filter compaction over Records (.y uniform in [0, 2^31), ~50 % selected).
    n = args.n_filter
    x = data.gen_uniform(n, seed, 0, 0, 0)
    y = data.gen_uniform(n, seed, 1, 0, 1 << 31)
    rec = DeviceBuffer(n * 8)
    qo._lib.call("qo_pack_records", x.ptr, y.ptr, n, rec.ptr, None)
    qo._lib.call("qo_device_sync")  # generated on the per-thread stream; timed on st
    x.free()
    out = DeviceBuffer(n * 8)
    cnt = DeviceBuffer(8)
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_filter_records", rec.ptr, n, 0, (1 << 30) - 1, out.ptr, cnt.ptr, st.handle))
    m = int(cnt.to_numpy(np.uint64, 1)[0])
    row("filter_records (map + scan + stable compaction)", n, ms, 8 * n + 8 * m, "8*N + 8*selected")
    out.free()
    rec.free()
    # Exclusive scan of int32.
    sc = DeviceBuffer(n * 4)
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_scan_exclusive_i32", y.ptr, sc.ptr, n, st.handle))
    row("scan_exclusive_i32", n, ms, 8 * n, "4*N read + 4*N written")
    sc.free()
    y.free()
    # Record sort and the reference's sort-based group-by pipeline.
    ns = args.n_sort
    G = 1 << 20
    x = data.gen_uniform(ns, seed, 2, 0, 1 << 16)
    y = data.gen_uniform(ns, seed, 3, 0, G)
    src = DeviceBuffer(ns * 8)
    qo._lib.call("qo_pack_records", x.ptr, y.ptr, ns, src.ptr, None)
    qo._lib.call("qo_device_sync")  # generated on the per-thread stream; timed on st
    x.free()
    y.free()
    work = DeviceBuffer(ns * 8)

    def sort_once():
        qo._lib.call("qo_copy", work.ptr, src.ptr, ns * 8, st.handle)
        qo._lib.call("qo_sort_records", work.ptr, ns, st.handle)

    ms_copy = _time_op(st, args.steps, args.warmup,
                       lambda: qo._lib.call("qo_copy", work.ptr, src.ptr, ns * 8, st.handle))
    ms = _time_op(st, args.steps, args.warmup, sort_once) - ms_copy
    row("sort_records (LSD radix on .y, 4 passes; the copy restoring the input is subtracted)", ns, ms,
        16 * ns, "8*N read + 8*N written")
    sp = DeviceBuffer(ns * 4)
    num = DeviceBuffer(8)
    aggs = [DeviceBuffer(G * 4) for _ in range(4)]

    def groupby_once():
        sort_once()
        qo._lib.call("qo_group_startpos", work.ptr, ns, sp.ptr, num.ptr, st.handle)
        for op in range(4):
            qo._lib.call("qo_aggregate_groups", work.ptr, ns, sp.ptr, G, op, aggs[op].ptr, st.handle)

    groupby_once()
    st.sync()
    g = int(num.to_numpy(np.uint64, 1)[0])
    if g != G:
        raise SystemExit(f"sort-based group-by found {g} groups, expected {G}")
    ms = _time_op(st, args.steps, args.warmup, groupby_once) - ms_copy
    row("sort-based group-by (sort + startPos + SUM/MAX/MIN/AVG), G = 2^20", ns, ms, 8 * ns + 28 * G,
        "8*N + 28*G")
    for b in aggs + [sp, num, work, src]:
        b.free()
    # Split, gather, scatter, projection, reduce and Record -> columns over
    # n_sort Records (.x = row id, .y uniform 32-bit).
    x = data.gen_perm(ns, seed + 1)
    y = data.gen_uniform(ns, seed, 4, 0, 0)
    src = DeviceBuffer(ns * 8)
    qo._lib.call("qo_pack_records", x.ptr, y.ptr, ns, src.ptr, None)
    # The inputs are generated on the per-thread stream and the operators timed
    # on st: wait for them here.  (The split reads its input twice, histogram
    # then placement; an input still being written between the two reads
    # gives placement offsets that do not match the rows placed.)
    qo._lib.call("qo_device_sync")
    dst = DeviceBuffer(ns * 8)
    bound = DeviceBuffer(65536 * 8)
    for parts, note in ((32, "one radix pass on the partition id"), (1000, "two radix passes on the partition id")):
        ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
            "qo_split_records", src.ptr, ns, parts, dst.ptr, bound.ptr, st.handle))
        row(f"split_records, {parts} partitions ({note})", ns, ms, 16 * ns, "8*N read + 8*N written")
    loc = x  # a permutation of [0, N): every gather / scatter target is distinct
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_gather_records", src.ptr, ns, loc.ptr, ns, dst.ptr, st.handle))
    row("gather_records (random permutation)", ns, ms, 20 * ns, "4*N loc + 8*N gathered + 8*N written")
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_scatter_records", src.ptr, ns, loc.ptr, dst.ptr, st.handle))
    row("scatter_records (random permutation)", ns, ms, 20 * ns, "8*N read + 4*N loc + 8*N scattered")
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_project_records", src.ptr, ns, src.ptr, ns, dst.ptr, st.handle))
    row("project_records (row ids a random permutation)", ns, ms, 20 * ns, "8*N proj + 4*N .y gathered + 8*N written")
    res = DeviceBuffer(4)
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_reduce_records", src.ptr, ns, 0, res.ptr, st.handle))
    row("reduce_records SUM (no grouping)", ns, ms, 8 * ns, "8*N read")
    cx, cy = DeviceBuffer(ns * 4), DeviceBuffer(ns * 4)
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_unpack_records", src.ptr, ns, cx.ptr, cy.ptr, st.handle))
    row("unpack_records (Records -> x, y columns)", ns, ms, 16 * ns, "8*N read + 8*N written")
    ms = _time_op(st, args.steps, args.warmup, lambda: qo._lib.call(
        "qo_pack_records", cx.ptr, cy.ptr, ns, dst.ptr, st.handle))
    row("pack_records (x, y columns -> Records)", ns, ms, 16 * ns, "8*N read + 8*N written")
    for b in (x, y, src, dst, bound, res, cx, cy):
        b.free()
    # Filter + join + sort (hj_filter_sort) on the join workload.
    nr, nsj = args.nr, args.ns
    t = data.JoinTables(nr, nsj, SEED_BASE + 2)
    tb = lib.qo_join_table_bytes(nr)
    table = DeviceBuffer(tb)
    outs = [DeviceBuffer(nsj * 4) for _ in range(3)]
    mm = C.c_uint64()
    ms = _time_op(st, max(1, args.steps // 2), 1, lambda: qo._lib.call(
        "qo_hj_filter_sort", table.ptr, tb, t.rkey.ptr, t.rpay.ptr, nr, t.skey.ptr, t.spay.ptr, t.sfilt.ptr, nsj,
        INT_MIN, 49, outs[0].ptr, outs[1].ptr, outs[2].ptr, nsj, C.byref(mm), st.handle))
    mj = mm.value
    row("hj_filter_sort (build, filtered probe, sort by (key, s_payload))", nsj, ms,
        16 * nr + 12 * nsj + 12 * mj + 24 * mj, "16*|R| + 12*|S| + 12*M + 24*M")
    for b in outs + [table]:
        b.free()
    t.free()
    worst = min(rows, key=lambda r: r["frac_hbm"])
    return {"metric": METRIC, "value": worst["GBps_alg"] * 1e9, "unit": "algorithmic bytes/s (slowest op)",
            "ms_per_step": worst["ms"], "dtype": "int32",
            "config": {"workload": "operator set of the reference path at scale", "ops": rows},
            "roofline": {"bound": "hbm", "kernel": worst["op"], "achieved": worst["GBps_alg"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": worst["frac_hbm"], "traffic": None,
                         "avg_launch_ms": worst["ms"]}}


def run_plan(args, qo, lib) -> dict:
    """The cost-model planner on filter -> group-by -> aggregate queries over
    n_sort Records: for every cover of each query's steps, the planner's
    estimate beside the measured time of qo_plan_execute (HIP events around the
    call on its stream; the call synchronises between segments), and which
    cover the planner picks.  With --plan-calibrate the primitive costs are
    measured on this device first."""
    import ctypes as C
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd import planner as P
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Stream

    n = args.n_sort
    st = Stream()
    if args.plan_calibrate:
        cls = min(26, max(16, n.bit_length() - 1))  # the workload's size class (at most 2^26 rows)
        log(f"[bench] plan: calibrating the primitive costs (2^{cls} rows)")
        P.calibrate(cls, 5)
    costs = {name: {"fixed_us": f, "gbps": g} for name, (f, g) in zip(P.COST_NAMES, P.cost_table())}
    # Implementation sets that force each cover (a query takes those that cover it).
    masks = [(1 << P.IMPL_FILTER) | (1 << P.IMPL_SORT_GROUP_AGG), (1 << P.IMPL_FILTER) | (1 << P.IMPL_HASH_GROUP_AGG),
             1 << P.IMPL_FILTER_HASH_GROUP_AGG]
    queries = []
    for G in (1 << 10, 1 << 20):
        queries.append((f"filter (half) -> group by .y (G = {G}) -> SUM(.x)", G,
                        P.Query(filter=True, lo=0, hi=G // 2 - 1, group=True, op=0, selectivity=0.5, groups=G)))
    queries.append(("group by .y (G = 2^20) -> MAX(.x)", 1 << 20,
                    P.Query(group=True, op=1, groups=1 << 20)))
    rows = []
    for name, G, q in queries:
        x = data.gen_uniform(n, SEED_BASE + 9, 0, 0, 0)
        y = data.gen_uniform(n, SEED_BASE + 9, 1, 0, G)
        rec = DeviceBuffer(n * 8)
        qo._lib.call("qo_pack_records", x.ptr, y.ptr, n, rec.ptr, None)
        qo._lib.call("qo_device_sync")
        x.free()
        y.free()
        keys, aggs, cnt = DeviceBuffer(G * 4), DeviceBuffer(G * 4), DeviceBuffer(8)
        qc = q.c()
        choice = P.plan(q, n)
        covers = []
        for mask in masks:
            try:
                pl = P.plan(q, n, mask)
            except qo._lib.QoError:
                continue
            if any(c["impls"] == pl.impls for c in covers):
                continue  # the same cover under a wider implementation set
            label = " -> ".join(P.IMPL_NAMES[i] for i in pl.impls)
            ms = _time_op(st, max(1, args.steps // 2), 1, lambda: qo._lib.call(
                "qo_plan_execute", C.byref(pl.raw), C.byref(qc), rec.ptr, n, G, None, keys.ptr, aggs.ptr, cnt.ptr,
                st.handle))
            g = int(cnt.to_numpy(np.uint64, 1)[0])
            covers.append({"cover": label, "impls": pl.impls, "est_ms": pl.est_ms, "measured_ms": ms, "groups": g,
                           "chosen": pl.impls == choice.impls})
            log(f"[bench] plan: {name}: {label}: est {pl.est_ms:.3f} ms, measured {ms:.3f} ms")
        best = min(covers, key=lambda c: c["measured_ms"])
        chosen = next(c for c in covers if c["chosen"])
        rows.append({"query": name, "n": n, "covers": covers, "planner_choice": chosen["cover"],
                     "fastest_measured": best["cover"],
                     "choice_over_fastest": chosen["measured_ms"] / best["measured_ms"]})
        for b in (rec, keys, aggs, cnt):
            b.free()
    worst = max(rows, key=lambda r: r["choice_over_fastest"])
    return {"metric": METRIC, "value": 1.0 / worst["choice_over_fastest"],
            "unit": "fastest measured cover time / planner-chosen cover time (worst query)", "ms_per_step": None,
            "dtype": "int32", "config": {"workload": "cost-model planner, filter -> group-by -> aggregate",
                                         "costs": costs, "queries": rows}}


def run_copy(args, qo, lib) -> dict:
    """Plain device copy of a known size: the achievable-HBM reference and the
    calibration kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Event, Stream

    nbytes = args.copy_bytes
    src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    qo._lib.call("qo_fill_u32", src.ptr, 7, nbytes // 4, None)
    st = Stream()
    fn = "qo_copy_nt" if args.copy_nt else "qo_copy"
    for _ in range(args.warmup):
        qo._lib.call(fn, dst.ptr, src.ptr, nbytes, st.handle)
    st.sync()
    evs = [(Event(), Event()) for _ in range(args.steps)]
    for a, b in evs:
        a.record(st)
        qo._lib.call(fn, dst.ptr, src.ptr, nbytes, st.handle)
        b.record(st)
    st.sync()
    ms = float(np.mean([a.elapsed_ms(b) for a, b in evs]))
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    src.free()
    dst.free()
    return {"metric": METRIC, "value": gbs, "unit": "GB/s (read+write)", "ms_per_step": ms, "dtype": "u8",
            "config": {"workload": "device copy (k_copy16)", "bytes": nbytes},
            "roofline": {"bound": "hbm", "kernel": "k_copy16", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": None, "alg_bytes_per_launch": 2 * nbytes,
                         "avg_launch_ms": ms}}


def _mk_batch_queries(args, qo, seed):
    """Config-4 style queries: |R| log-uniform in [4K, 4M] with unique 31-bit keys,
    |S| = |R| * U[4, 16], fk uniform over R, filter U[0,1000) < 1000*sigma,
    sigma ~ U[0.05, 0.5].  All columns generated in HBM."""
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd.device import DeviceBuffer

    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    qs = []
    for i in range(args.queries):
        u = rng.random()
        nr = int(2 ** (12 + 10 * u * args.batch_scale))
        ns = int(nr * (4 + 12 * rng.random()))
        sigma = 0.05 + 0.45 * rng.random()
        fhi = int(1000 * sigma) - 1
        qseed = seed + 1000 + i
        q = {"nr": nr, "ns": ns, "flo": INT_MIN, "fhi": fhi,
             "rkey": data.gen_bij(nr, 31, qseed), "rpay": data.gen_uniform(nr, qseed, 1, 0, 0),
             "skey": data.gen_fk(ns, qseed, 2, nr, 31, qseed), "spay": data.gen_uniform(ns, qseed, 3, 0, 0),
             "sfilt": data.gen_uniform(ns, qseed, 4, 0, 1000)}
        q["cap"] = int(ns * (fhi + 1) / 1000 * 1.1) + 4096
        for c in ("okey", "orpay", "ospay"):
            q[c] = DeviceBuffer(q["cap"] * 4)
        qs.append(q)
    qo._lib.call("qo_device_sync")
    return qs


def _digest(qo, q, m, buf):
    """(checksum, key sum, order violations) of a query's first m output rows."""
    qo._lib.call("qo_join_rows_digest", q["okey"].ptr, q["orpay"].ptr, q["ospay"].ptr, int(m), None, buf.ptr, None)
    return tuple(int(x) for x in buf.to_numpy(np.uint64, 3))


def run_batched(args, qo, lib, scale=None, label="configured sizes") -> dict:
    """Batched hj_filter_sort: all queries in one graph launch vs the plain path
    (qo_hj_filter_sort per query, synchronous like the reference's host loop).
    Every query's batched output is checked against its plain output (match
    count and an order-sensitive device digest of the three columns, which also
    counts rows out of (key, s_payload) order)."""
    import ctypes as C

    from query_on_opencl_fpga_amd.device import DeviceBuffer, Event, Stream
    from query_on_opencl_fpga_amd.operators import JoinBatch

    seed = SEED_BASE + 4
    if scale is not None:
        args = argparse.Namespace(**dict(vars(args), batch_scale=scale))
    log(f"[bench] batched ({label}): generating {args.queries} queries")
    qs = _mk_batch_queries(args, qo, seed)
    nq = len(qs)
    tot_r = sum(q["nr"] for q in qs)
    tot_s = sum(q["ns"] for q in qs)
    log(f"[bench] batched: |R| total {tot_r}, |S| total {tot_s}; creating batch")
    b = JoinBatch(qs)
    st = Stream()
    for _ in range(args.warmup):
        b.launch(st)
    st.sync()
    evs = [(Event(), Event()) for _ in range(args.steps)]
    for e0, e1 in evs:
        e0.record(st)
        b.launch(st)
        e1.record(st)
    st.sync()
    launch_ms = np.array([e0.elapsed_ms(e1) for e0, e1 in evs])
    t_batch = float(launch_ms.mean())
    counts_b = [int(c) for c in b.counts()]
    dig = DeviceBuffer(24)
    dig_b = [_digest(qo, q, m, dig) for q, m in zip(qs, counts_b)]
    b.destroy()
    log(f"[bench] batched: {t_batch:.2f} ms per batch; plain path")
    maxr = max(q["nr"] for q in qs)
    table = DeviceBuffer(lib.qo_join_table_bytes(maxr))
    m = C.c_uint64()

    def plain_all(record):
        lat = []
        t0 = time.perf_counter()
        for q in qs:
            t1 = time.perf_counter()
            qo._lib.call("qo_hj_filter_sort", table.ptr, lib.qo_join_table_bytes(q["nr"]), q["rkey"].ptr,
                         q["rpay"].ptr, q["nr"], q["skey"].ptr, q["spay"].ptr, q["sfilt"].ptr, q["ns"], q["flo"],
                         q["fhi"], q["okey"].ptr, q["orpay"].ptr, q["ospay"].ptr, q["cap"], C.byref(m), st.handle)
            lat.append((time.perf_counter() - t1) * 1e3)
            if record:
                counts_p.append(m.value)
                dig_p.append(_digest(qo, q, m.value, dig))
        return (time.perf_counter() - t0) * 1e3, np.array(lat)

    counts_p, dig_p = [], []
    plain_all(False)  # warm-up (workspaces grow to the largest query)
    t_plain, lat = plain_all(False)
    plain_all(True)   # the checked run (digests outside the timed run)
    bad = [i for i in range(nq) if counts_b[i] != counts_p[i] or dig_b[i] != dig_p[i] or dig_b[i][2] != 0]
    if bad:
        i = bad[0]
        raise SystemExit(f"batched and plain results differ for {len(bad)} queries; first {i}: "
                         f"batched {counts_b[i]} {dig_b[i]} plain {counts_p[i]} {dig_p[i]}")
    table.free()
    dig.free()
    alg = sum(16 * q["nr"] + 12 * q["ns"] + 36 * c for q, c in zip(qs, counts_b))
    gbs = alg / (t_batch * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": nq / (t_batch * 1e-3), "unit": "queries/s (batched)",
        "ms_per_step": t_batch, "dtype": "int32",
        "config": {"workload": f"batched hj_filter_sort ({label}): filter S, build R, probe, sort by (key, s_payload)",
                   "queries": nq, "R_log2_range": [12, 12 + 10 * args.batch_scale], "R_total": tot_r,
                   "S_total": tot_s, "matches_total": int(sum(counts_p)),
                   "checked": "every query: batched count and device digest (checksum, key sum, order) == plain"},
        "batched": {"queries_per_s": nq / (t_batch * 1e-3), "launches": args.steps,
                    "p50_ms": float(np.percentile(launch_ms, 50)), "p99_ms": float(np.percentile(launch_ms, 99)),
                    "latency": "a query's latency is its batch's launch time (HIP events around each graph launch)"},
        "plain": {"queries_per_s": nq / (t_plain * 1e-3), "p50_ms": float(np.percentile(lat, 50)),
                  "p99_ms": float(np.percentile(lat, 99)), "total_ms": t_plain,
                  "latency": "each query alone, host wall time of its synchronous qo_hj_filter_sort call",
                  "common_arrival": {
                      "p50_ms": float(np.percentile(np.cumsum(lat), 50)),
                      "p99_ms": float(np.percentile(np.cumsum(lat), 99)),
                      "latency": "all queries arrive together and run one after the other: a query's "
                                 "latency is the time until it finishes (the batched latency's counterpart)"}},
        "speedup_batched_vs_plain": t_plain / t_batch,
        "roofline": {"bound": "hbm", "kernel": "one hipGraph launch of the whole batch (all stages)",
                     "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_launch": alg,
                     "alg_bytes_formula": "sum over queries of 16*|R| + 12*|S| + 12*M + 24*M", "avg_launch_ms": t_batch},
    }
    if not args.no_cpu_baseline:
        k = min(nq, args.cpu_batch_queries)
        log(f"[bench] batched: cpu baseline on {k} queries")
        res["cpu_baseline"] = cpu_baseline_batch(qs[:k])
    for q in qs:
        for v in q.values():
            if hasattr(v, "free"):
                v.free()
    return res


def _serve(qo, tables, schedule, plain: bool, args) -> dict:
    """Queues `schedule` (list of (t_ms, request tuple)) with its arrival times on
    the server clock, serves it and returns per-request latencies (end - arrival)."""
    from query_on_opencl_fpga_amd.serving import Server

    sv = Server(budget_bytes=args.budget_gb * 1e9, inflight=1, batch=args.batch_max, plain=plain)
    ids = [sv.add_table(t["n"], key=t["key"], fk=t["fk"], grp=t["grp"], val=t["val"], filt=t["filt"],
                        group_capacity=t["G"]) for t in tables]
    # Warm-up: one of each kind (workspaces and pools reach steady state).
    warm = [(k, ids[i % len(ids)], ids[-1], INT_MIN, 499) for i, k in enumerate((0, 1, 2))]
    for r in warm:
        sv.submit(*r)
    sv.run()
    sv.poll(timeout_ms=0)
    base = sv.now_ms() + 2.0
    for t_ms, (kind, a, b, lo, hi) in schedule:
        sv.submit_at(kind, ids[a], ids[b] if b >= 0 else -1, lo, hi, arrival_ms=base + t_ms)
    sv.run()
    res = sv.poll(max_results=len(schedule) + 16, timeout_ms=0)
    sv.destroy()
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise SystemExit(f"{len(bad)} served requests failed, first: {bad[0]}")
    lat = np.array([r["end_ms"] - r["submit_ms"] for r in res])
    span = max(r["end_ms"] for r in res) - base
    return {"queries_per_s": len(res) / (span * 1e-3), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "span_ms": span,
            "checks": {r["id"]: (r["count"], r["sum"], r["checksum"]) for r in res}}


def run_served(args, qo, lib) -> dict:
    """Served mixed stream: Poisson arrivals of filter (40%), group-by (30%) and
    join+sort (30%) over a pool of HBM-resident tables (sizes log-uniform in
    [2^16, 2^26], group cardinality log-uniform in [16, 2^20], Zipf z=0.8),
    continuous batching vs the plain per-query path on the same arrivals."""
    from query_on_opencl_fpga_amd import data

    seed = SEED_BASE + 5
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    tables = []
    log(f"[bench] served: generating {args.tables} tables")
    for i in range(args.tables):
        n = int(2 ** rng.uniform(16, 16 + 10 * args.batch_scale))
        G = int(2 ** rng.uniform(4, 20))
        s = seed + 100 + i
        tables.append({"n": n, "G": G, "key": data.gen_bij(n, 31, seed), "fk": data.gen_fk(n, s, 1, n, 31, seed),
                       "grp": data.gen_zipf(n, G, 0.8, s, 2, s + 7), "val": data.gen_uniform(n, s, 3, 0, 1 << 16),
                       "filt": data.gen_uniform(n, s, 4, 0, 1000)})
    qo._lib.call("qo_device_sync")
    sched, t = [], 0.0
    for _ in range(args.requests):
        t += rng.exponential(1000.0 / args.qps)
        u = rng.random()
        a = int(rng.integers(len(tables)))
        hi = int(rng.integers(50, 500))
        if u < 0.4:
            req = (0, a, -1, INT_MIN, hi)
        elif u < 0.7:
            req = (1, a, -1, INT_MIN, hi)
        else:
            b = int(rng.integers(len(tables)))
            lo_t, hi_t = sorted((a, b), key=lambda k: tables[k]["n"])
            req = (2, lo_t, hi_t, INT_MIN, hi)
        sched.append((t, req))
    log(f"[bench] served: {args.requests} requests at {args.qps} qps (served mode)")
    served = _serve(qo, tables, sched, False, args)
    log("[bench] served: plain mode")
    plain = _serve(qo, tables, sched, True, args)
    same = sorted(served.pop("checks").values()) == sorted(plain.pop("checks").values())
    if not same:
        raise SystemExit("served and plain results differ")
    for tb in tables:
        for k in ("key", "fk", "grp", "val", "filt"):
            tb[k].free()
    return {
        "metric": METRIC, "value": served["queries_per_s"], "unit": "queries/s (served)",
        "ms_per_step": served["span_ms"], "dtype": "int32",
        "config": {"workload": "served mixed stream (filter 40%, group-by 30%, join+sort 30%), Poisson arrivals",
                   "offered_qps": args.qps, "requests": args.requests, "tables": args.tables,
                   "table_rows_total": int(sum(tb["n"] for tb in tables)), "budget_gb": args.budget_gb,
                   "inflight_batches": 1, "max_batch": args.batch_max},
        "served": served, "plain": plain,
        "throughput_ratio_served_vs_plain": served["queries_per_s"] / plain["queries_per_s"],
        "p99_ratio_plain_vs_served": plain["p99_ms"] / served["p99_ms"],
        "results_identical": same,
    }


def run_groupby(args, qo, lib) -> dict:
    """Hash group-by (COUNT, SUM, MIN, MAX per group, sorted by key) at N tuples,
    key Zipf(z) over G groups, for G in {2^10, 2^20} and z in {0, 0.5, 1}; every
    case with its roofline (8*N + 28*G algorithmic bytes over the HIP-event time
    of one qo_groupby_agg call) and a CPU baseline on a 2^24-tuple sample."""
    from query_on_opencl_fpga_amd import data
    from query_on_opencl_fpga_amd.device import DeviceBuffer, Event, Stream

    n = args.n_groupby
    seed = SEED_BASE + 3
    rows = []
    st = Stream()
    for G in (1 << 10, 1 << 20):
        for z in (0.0, 0.5, 1.0):
            log(f"[bench] groupby: n={n} G={G} z={z}")
            t = data.GroupbyTables(n, G, z, seed)
            outs = [DeviceBuffer(G * s) for s in (4, 8, 8, 4, 4)]
            num = DeviceBuffer(8)

            def step():
                qo._lib.call("qo_groupby_agg", t.key.ptr, t.val.ptr, n, G, *[b.ptr for b in outs], num.ptr,
                             st.handle)

            ms = _time_op(st, args.steps, args.warmup, step)
            g = int(num.to_numpy(np.uint64, 1)[0])
            cnt = outs[1].to_numpy(np.int64, g)
            keys = outs[0].to_numpy(np.int32, g)
            if int(cnt.sum()) != n or g > G or (g > 1 and not np.all(np.diff(keys) > 0)):
                raise SystemExit(f"group-by G={G} z={z}: {g} groups, count sum {int(cnt.sum())} (expected {n})")
            alg = 8 * n + 28 * G
            gbs = alg / (ms * 1e-3) / 1e9
            kernel = ("qo_groupby_agg, partitioned (k_part_hist + scan + k_part_place + k_groupby_chunks + "
                      "collect/sort/gather)" if G > 1024 else
                      "qo_groupby_agg, direct (k_groupby_hash + collect/sort/gather)")
            row = {"G": G, "z": z, "groups": g, "ms": ms, "tuples_per_s": n / (ms * 1e-3),
                   "roofline": {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                                "alg_bytes_per_launch": alg, "alg_bytes_formula": "8*N + 28*G",
                                "avg_launch_ms": ms}}
            for b in outs + [num]:
                b.free()
            if not args.no_cpu_baseline:
                sn = min(n, args.cpu_groupby_sample)
                k, v = t.key.to_numpy(np.int32, sn), t.val.to_numpy(np.int32, sn)
                row["cpu_baseline"] = cpu_baseline_groupby(k, v, f"first {sn} of the case's {n} tuples")
                del k, v
            t.free()
            rows.append(row)
    worst = min(rows, key=lambda r: r["tuples_per_s"])
    return {"metric": METRIC, "value": worst["tuples_per_s"], "unit": "tuples/s (slowest of 6 cases)",
            "ms_per_step": worst["ms"], "dtype": "int32",
            "config": {"workload": "hash group-by (COUNT, SUM, MIN, MAX), key Zipf(z) over G groups", "tuples": n,
                       "cases": rows},
            "roofline": dict(worst["roofline"], case=f"G={worst['G']} z={worst['z']}"),
            "cpu_baseline": worst.get("cpu_baseline")}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="join",
                    choices=["join", "filter_agg", "groupby", "batched", "served", "copy", "ops", "plan"])
    ap.add_argument("--plan-calibrate", action="store_true",
                    help="plan config: measure the planner's primitive costs on this device first")
    ap.add_argument("--load", type=int, default=0, help="join table load factor in percent (default: tuned)")
    ap.add_argument("--join-mode", choices=["auto", "partitioned", "direct", "lds"], default="auto",
                    help="join build/probe strategy (default: partitioned once the table outgrows one L2)")
    ap.add_argument("--slice-kib", type=int, default=0, help="partitioned join: table slice per partition")
    ap.add_argument("--no-key-range", action="store_true",
                    help="join: always build the hash-mixed table (no key-range table)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copy-nt", action="store_true", help="copy with non-temporal loads")
    ap.add_argument("--queries", type=int, default=512)
    ap.add_argument("--qps", type=float, default=1000.0, help="served mode: offered load")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--tables", type=int, default=16)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--batch-max", type=int, default=64)
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="fraction of the log2 size range for batched queries (1.0 = [4K, 4M])")
    ap.add_argument("--nr", type=int, default=1 << 24)
    ap.add_argument("--ns", type=int, default=1 << 28)
    ap.add_argument("--n-filter", type=int, default=1 << 28)
    ap.add_argument("--n-groupby", type=int, default=1 << 28)
    ap.add_argument("--n-sort", type=int, default=1 << 26)
    ap.add_argument("--cpu-groupby-sample", type=int, default=1 << 24,
                    help="group-by CPU baseline: tuples of each case's sample")
    ap.add_argument("--cpu-batch-queries", type=int, default=16,
                    help="batched CPU baseline: queries of the sample (the first ones)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="default config: measure only the headline join (no per-config sub-objects)")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the roofline kernel from rocprofv3 --pmc, if measured")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the operators shard by "
              "range/hash across GPUs by design, but this build measures one GPU)", file=sys.stderr)
        return 2
    import query_on_opencl_fpga_amd as qo

    lib = qo._lib.load()
    qo._lib.call("qo_init", 0)
    info = qo.device.device_info()
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['cu_count']} LDS/block={info['lds_bytes_per_block']}")
    if args.join_mode != "auto" or args.slice_kib:
        import ctypes as C
        mode = {"auto": 0, "partitioned": 1, "direct": 2, "lds": 3}[args.join_mode]
        for cls in range(40):
            b, i, p = C.c_int(), C.c_int(), C.c_int()
            qo._lib.call("qo_tune_get", 5, cls, C.byref(b), C.byref(i), C.byref(p))
            qo._lib.call("qo_tune_set", 5, cls, b.value, args.slice_kib or i.value, mode)
    if args.no_key_range:
        for cls in range(40):
            qo._lib.call("qo_tune_set", 7, cls, 256, 256, 2)
    run = {"join": run_join, "filter_agg": run_filter_agg, "groupby": run_groupby, "copy": run_copy,
           "batched": run_batched, "served": run_served, "ops": run_ops, "plan": run_plan}
    if args.config not in run:
        print(f"config {args.config} is not available", file=sys.stderr)
        return 2
    res = run[args.config](args, qo, lib)
    if args.config == "join" and not args.headline_only:
        # The other configurations, each with its roofline and CPU baseline
        # (fewer timed steps: every one is a steady-state kernel measurement).
        sub = argparse.Namespace(**dict(vars(args), steps=min(args.steps, 5), warmup=min(args.warmup, 2)))
        configs = {}
        for name, fn in (("filter_agg_2e28", lambda: run_filter_agg(sub, qo, lib)),
                         ("groupby_2e28", lambda: run_groupby(sub, qo, lib)),
                         ("batched_512q", lambda: run_batched(sub, qo, lib, 1.0, "|R| log-uniform in [4K, 4M]")),
                         ("batched_512q_small", lambda: run_batched(sub, qo, lib, 0.4,
                                                                    "small queries, |R| log-uniform in [4K, 64K]"))):
            r = fn()
            r.pop("metric", None)
            r["steps"], r["warmup"] = sub.steps, sub.warmup
            configs[name] = r
            qo._lib.call("qo_release_workspaces")
        res["configs"] = configs
    line = {
        "metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"),
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
        "data": "synthetic (deterministic splitmix64 columns generated in HBM)",
    }
    line.update(res)
    line["device"] = {"name": info["name"], "arch": info["arch"], "cu_count": info["cu_count"],
                      "host_cores": os.cpu_count()}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
