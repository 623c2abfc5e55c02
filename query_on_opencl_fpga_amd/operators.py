"""Host-side operator interface mirroring the reference's entry points.

Records are numpy ``uint32`` arrays of shape ``(n, 2)``: column 0 is ``.x``
(value / row id), column 1 is ``.y`` (key), the reference's ``Record = uint2``
(host/host/common.h:27-28).  Functions named like the reference
(``CL_RangeSelection``, ``CL_RadixSort``, ``CL_GroupBy``,
``CL_agg_*_afterGroupBy``, ``buildHashTable`` / ``probeHashTable``) call the
reference-named C entry points of libqo; the lower-case functions call the
native column API.  Every call runs on the GPU through libqo; nothing here
computes a result on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceBuffer

REDUCE_SUM, REDUCE_MAX, REDUCE_MIN, REDUCE_AVERAGE = 0, 1, 2, 3
TEST_MAX = 1 << 30


def _records(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("records must have shape (n, 2) = (x, y)")
    return a


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _free_c(ptr) -> None:
    C.CDLL(None).free(C.c_void_p(ptr))


# ---------------------------------------------------------------------------
# Reference-named host-pointer entry points (host/host/Selection.cpp:19,
# Sort.cpp:34, GroupBy.cpp:7, AggAfterGroupBy.cpp:30-170).
# ---------------------------------------------------------------------------
def CL_RangeSelection(h_Rin, rangeSmallKey: int, rangeLargeKey: int, numThreadPB: int = 256,
                      numBlock: int = 128) -> np.ndarray:
    """Records with rangeSmallKey <= (int)y <= rangeLargeKey, in input order."""
    rin = _records(h_Rin)
    lib = _lib.load()
    out = C.c_void_p()
    n = lib.CL_RangeSelection(rin.ctypes.data, rin.shape[0], int(rangeSmallKey), int(rangeLargeKey),
                              C.byref(out), numThreadPB, numBlock, 0)
    if n < 0:
        raise _lib.QoError(n, "CL_RangeSelection", _lib.last_error())
    res = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint32)), shape=(max(n, 1) * 2,))[: 2 * n]
    res = res.reshape(n, 2).copy()
    _free_c(out.value)
    return res


def CL_PointSelection(h_Rin, matchingKeyValue: int, numThreadPB: int = 256, numBlock: int = 128) -> np.ndarray:
    """Records with (int)y == matchingKeyValue, in input order (Selection.cpp:65)."""
    rin = _records(h_Rin)
    lib = _lib.load()
    out = C.c_void_p()
    n = lib.CL_PointSelection(rin.ctypes.data, rin.shape[0], int(matchingKeyValue), C.byref(out), numThreadPB,
                              numBlock, 0)
    if n < 0:
        raise _lib.QoError(n, "CL_PointSelection", _lib.last_error())
    res = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint32)), shape=(max(n, 1) * 2,))[: 2 * n]
    res = res.reshape(n, 2).copy()
    _free_c(out.value)
    return res


def _mem_call(name: str, lib, *args) -> None:
    _clear_error()
    getattr(lib, name)(*args)
    _raise_if_error(name)


def splitImpl(h_R, numPart: int, numThreadPB: int = 128, numBlock: int = 128):
    """(split records, bounds): records grouped by y % numPart, input order kept
    within a partition; bounds[p] = (start, end) (testSplit.cpp:156-209)."""
    r = _records(h_R)
    n = r.shape[0]
    lib = _lib.load()
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer(max(n, 1) * 8) as d_s, \
            DeviceBuffer(max(int(numPart), 1) * 8) as d_b:
        m = [lib.qo_mem_wrap(b.ptr, b.nbytes) for b in (d_r, d_s, d_b)]
        try:
            _mem_call("splitImpl", lib, m[0], n, int(numPart), m[1], m[2], numThreadPB, numBlock, 4,
                      None, None, None, None, None, 0)
        finally:
            for x in m:
                lib.qo_mem_free(x)
        return d_s.to_numpy(np.uint32, 2 * n).reshape(n, 2), d_b.to_numpy(np.uint32, 2 * int(numPart)).reshape(-1, 2)


def scatterImpl(h_R, loc, h_S=None, numThreadPB: int = 256, numBlock: int = 128) -> np.ndarray:
    """S[loc[i]] = R[i] (testScatter.cpp:61-81) over S (default zeros, len(R) records)."""
    r = _records(h_R)
    lc = _i32(loc)
    s = np.zeros_like(r) if h_S is None else _records(h_S)
    lib = _lib.load()
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer.from_numpy(lc) as d_l, DeviceBuffer.from_numpy(s) as d_s:
        m = [lib.qo_mem_wrap(b.ptr, b.nbytes) for b in (d_r, d_l, d_s)]
        try:
            _mem_call("scatterImpl", lib, m[0], r.shape[0], m[1], m[2], numThreadPB, numBlock,
                      None, None, None, None, None, 0)
        finally:
            for x in m:
                lib.qo_mem_free(x)
        return d_s.to_numpy(np.uint32, s.size).reshape(-1, 2)


def gatherImpl(h_R, loc, h_S=None, numThreadPB: int = 256, numBlock: int = 128) -> np.ndarray:
    """S[i] = R[loc[i]] (testScatter.cpp:105-130) over S (default zeros, len(loc) records)."""
    r = _records(h_R)
    lc = _i32(loc)
    s = np.zeros((lc.shape[0], 2), np.uint32) if h_S is None else _records(h_S)
    lib = _lib.load()
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer.from_numpy(lc) as d_l, DeviceBuffer.from_numpy(s) as d_s:
        m = [lib.qo_mem_wrap(b.ptr, b.nbytes) for b in (d_r, d_l, d_s)]
        try:
            _mem_call("gatherImpl", lib, m[0], r.shape[0], m[1], m[2], lc.shape[0], numThreadPB, numBlock,
                      None, None, None, None, None, 0)
        finally:
            for x in m:
                lib.qo_mem_free(x)
        return d_s.to_numpy(np.uint32, s.size).reshape(-1, 2)


def CL_RadixSort(h_Rin, numThread: int = 256, numBlock: int = 128) -> np.ndarray:
    """Records sorted ascending by y (uint32); equal keys keep input order."""
    rin = _records(h_Rin)
    out = np.empty_like(rin)
    lib = _lib.load()
    _clear_error()
    lib.CL_RadixSort(rin.ctypes.data, rin.shape[0], out.ctypes.data, numThread, numBlock, 0)
    _raise_if_error("CL_RadixSort")
    return out


def CL_GroupBy(h_Rin, numThread: int = 256, numBlock: int = 128):
    """(sorted records, startPos) as groupByImpl computes them."""
    rin = _records(h_Rin)
    out = np.empty_like(rin)
    sp = C.c_void_p()
    lib = _lib.load()
    g = lib.CL_GroupBy(rin.ctypes.data, rin.shape[0], out.ctypes.data, C.byref(sp), numThread, numBlock, 0)
    if g < 0:
        raise _lib.QoError(g, "CL_GroupBy", _lib.last_error())
    start = np.ctypeslib.as_array(C.cast(sp, C.POINTER(C.c_int32)), shape=(max(g, 1),))[:g].copy()
    _free_c(sp.value)
    return out, start


def _cl_agg(name: str, h_Rin, startPos, h_Ragg=None, numThread: int = 256) -> np.ndarray:
    rin = _records(h_Rin)
    sp = _i32(startPos)
    res = np.zeros(sp.shape[0], dtype=np.int32)
    ragg = None if h_Ragg is None else _records(h_Ragg)
    lib = _lib.load()
    _clear_error()
    getattr(lib, name)(rin.ctypes.data, rin.shape[0], sp.ctypes.data, sp.shape[0],
                       None if ragg is None else ragg.ctypes.data, res.ctypes.data, numThread, 0)
    _raise_if_error(name)
    return res


def CL_agg_max_afterGroupBy(h_Rin, startPos, h_Ragg=None, numThread: int = 256):
    return _cl_agg("CL_agg_max_afterGroupBy", h_Rin, startPos, h_Ragg, numThread)


def CL_agg_min_afterGroupBy(h_Rin, startPos, h_Ragg=None, numThread: int = 256):
    return _cl_agg("CL_agg_min_afterGroupBy", h_Rin, startPos, h_Ragg, numThread)


def CL_agg_sum_afterGroupBy(h_Rin, startPos, h_Ragg=None, numThread: int = 256):
    return _cl_agg("CL_agg_sum_afterGroupBy", h_Rin, startPos, h_Ragg, numThread)


def CL_agg_avg_afterGroupBy(h_Rin, startPos, h_Ragg=None, numThread: int = 256):
    return _cl_agg("CL_agg_avg_afterGroupBy", h_Rin, startPos, h_Ragg, numThread)


def scanImpl(values) -> np.ndarray:
    """Exclusive int32 prefix sum (host/host/testScan.cpp:278)."""
    v = _i32(values)
    n = v.shape[0]
    if n == 0:
        return v.copy()
    lib = _lib.load()
    with DeviceBuffer.from_numpy(v) as src, DeviceBuffer(v.nbytes) as dst:
        msrc, mdst = lib.qo_mem_wrap(src.ptr, v.nbytes), lib.qo_mem_wrap(dst.ptr, v.nbytes)
        sp = _lib.ScanPara()
        lib.initScan(n, C.byref(sp))
        _clear_error()
        lib.scanImpl(msrc, n, mdst, None, None, None, None, None, C.byref(sp), 0)
        lib.closeScan(C.byref(sp))
        lib.qo_mem_free(msrc)
        lib.qo_mem_free(mdst)
        _raise_if_error("scanImpl")
        return dst.to_numpy(np.int32, n)


def buildHashTable(rTable, B: int, bucketSize: int):
    """Bucket-chained build (hj_filter_sort_op.cl:27) from interleaved (key, val) pairs.

    Returns device buffers (table, counts) for probeHashTable."""
    r = np.ascontiguousarray(rTable, dtype=np.uint32).reshape(-1)
    size = r.shape[0] // 2
    lib = _lib.load()
    rbuf = DeviceBuffer.from_numpy(r)
    table = DeviceBuffer(B * bucketSize * 8)
    counts = DeviceBuffer(B * 4)
    rc = lib.buildHashTable(rbuf.ptr, table.ptr, size, B, bucketSize, counts.ptr)
    rbuf.free()
    if rc != 0:
        table.free()
        counts.free()
        raise _lib.QoError(rc, "buildHashTable", _lib.last_error())
    return table, counts


def probeHashTable(table: DeviceBuffer, counts: DeviceBuffer, sTable, B: int, bucketSize: int) -> np.ndarray:
    """Compacted (s_val, r_val) pairs of every match (hj_filter_sort_op.cl:64)."""
    s = np.ascontiguousarray(sTable, dtype=np.uint32).reshape(-1)
    size = s.shape[0] // 2
    lib = _lib.load()
    with DeviceBuffer.from_numpy(s) as sbuf, DeviceBuffer(max(size, 1) * 8) as mbuf:
        m = lib.probeHashTable(table.ptr, sbuf.ptr, mbuf.ptr, size, B, bucketSize, counts.ptr)
        if m < 0:
            raise _lib.QoError(m, "probeHashTable", _lib.last_error())
        return mbuf.to_numpy(np.uint32, 2 * m).reshape(m, 2)


# ---------------------------------------------------------------------------
# Native column API (include/qo.h), host arrays in and out.
# ---------------------------------------------------------------------------
def filter_agg(key, val, lo: int, hi: int):
    """(COUNT, SUM int64, SUM wrapped to int32) of val where lo <= key <= hi."""
    k, v = _i32(key), _i32(val)
    with DeviceBuffer.from_numpy(k) as dk, DeviceBuffer.from_numpy(v) as dv, DeviceBuffer(24) as out:
        _lib.call("qo_filter_agg", dk.ptr, dv.ptr, k.shape[0], int(lo), int(hi), out.ptr, None)
        r = out.to_numpy(np.int64, 3)
    return int(r[0]), int(r[1]), int(r[2])


def filter_records(h_Rin, lo: int, hi: int) -> np.ndarray:
    rin = _records(h_Rin)
    n = rin.shape[0]
    with DeviceBuffer.from_numpy(rin) as d_in, DeviceBuffer(max(n, 1) * 8) as d_out, DeviceBuffer(8) as cnt:
        _lib.call("qo_filter_records", d_in.ptr, n, int(lo), int(hi), d_out.ptr, cnt.ptr, None)
        m = int(cnt.to_numpy(np.uint64, 1)[0])
        return d_out.to_numpy(np.uint32, 2 * m).reshape(m, 2)


def groupby_agg(key, val, group_capacity: int):
    """Per key: (keys, COUNT, SUM, MIN, MAX) ascending by key."""
    k, v = _i32(key), _i32(val)
    g = int(group_capacity)
    gc = max(g, 1)
    bufs = [DeviceBuffer(gc * s) for s in (4, 8, 8, 4, 4)]
    with DeviceBuffer.from_numpy(k) as dk, DeviceBuffer.from_numpy(v) as dv, DeviceBuffer(8) as num:
        _lib.call("qo_groupby_agg", dk.ptr, dv.ptr, k.shape[0], g, *[b.ptr for b in bufs], num.ptr, None)
        ng = int(num.to_numpy(np.uint64, 1)[0])
    if ng > g:
        for b in bufs:
            b.free()
        raise _lib.QoError(3, "groupby_agg", f"{ng} groups exceed capacity {g}")
    out = tuple(b.to_numpy(dt, ng) for b, dt in zip(bufs, (np.int32, np.int64, np.int64, np.int32, np.int32)))
    for b in bufs:
        b.free()
    return out


def hash_join(rkey, rpay, skey, spay, sfilt=None, flo: int = -(1 << 31), fhi: int = (1 << 31) - 1,
              capacity: int | None = None):
    """(key, r_payload, s_payload) of every match (row order unspecified) and the match count."""
    rk, rp, sk, sp = _i32(rkey), _i32(rpay), _i32(skey), _i32(spay)
    nr, ns = rk.shape[0], sk.shape[0]
    cap = ns if capacity is None else int(capacity)
    lib = _lib.load()
    tb = lib.qo_join_table_bytes(nr)
    bufs = [DeviceBuffer.from_numpy(a) for a in (rk, rp, sk, sp)]
    fbuf = DeviceBuffer.from_numpy(_i32(sfilt)) if sfilt is not None else None
    outs = [DeviceBuffer(max(cap, 1) * 4) for _ in range(3)]
    with DeviceBuffer(tb) as table, DeviceBuffer(8) as cnt:
        _lib.call("qo_hash_build", table.ptr, tb, bufs[0].ptr, bufs[1].ptr, nr, None)
        _lib.call("qo_hash_probe", table.ptr, tb, bufs[2].ptr, bufs[3].ptr, fbuf.ptr if fbuf else None, ns,
                  int(flo), int(fhi), outs[0].ptr, outs[1].ptr, outs[2].ptr, cap, cnt.ptr, None)
        m = int(cnt.to_numpy(np.uint64, 1)[0])
    res = tuple(o.to_numpy(np.int32, min(m, cap)) for o in outs)
    for b in bufs + outs + ([fbuf] if fbuf else []):
        b.free()
    return res + (m,)


def hj_filter_sort(rkey, rpay, skey, spay, sfilt, flo: int, fhi: int, capacity: int | None = None):
    """Filter S, join with R, sort by (key, s_payload): returns (key, r_payload, s_payload)."""
    rk, rp, sk, sp, sf = _i32(rkey), _i32(rpay), _i32(skey), _i32(spay), _i32(sfilt)
    nr, ns = rk.shape[0], sk.shape[0]
    cap = ns if capacity is None else int(capacity)
    lib = _lib.load()
    tb = lib.qo_join_table_bytes(nr)
    bufs = [DeviceBuffer.from_numpy(a) for a in (rk, rp, sk, sp, sf)]
    outs = [DeviceBuffer(max(cap, 1) * 4) for _ in range(3)]
    m = C.c_uint64()
    with DeviceBuffer(tb) as table:
        _lib.call("qo_hj_filter_sort", table.ptr, tb, bufs[0].ptr, bufs[1].ptr, nr, bufs[2].ptr, bufs[3].ptr,
                  bufs[4].ptr, ns, int(flo), int(fhi), outs[0].ptr, outs[1].ptr, outs[2].ptr, cap,
                  C.byref(m), None)
    res = tuple(o.to_numpy(np.int32, m.value) for o in outs)
    for b in bufs + outs:
        b.free()
    return res


class JoinBatch:
    """A batch of filtered join+sort queries over device-resident columns,
    captured once and replayed with one graph launch (qo_batch_*).

    Each query is a dict of DeviceBuffers / ints: rkey, rpay, skey, spay, sfilt
    (optional), nr, ns, flo, fhi, okey, orpay, ospay, cap."""

    def __init__(self, queries):
        arr = (_lib.QoJoinQuery * len(queries))()
        for i, q in enumerate(queries):
            a = arr[i]
            a.rkey, a.rpay = q["rkey"].ptr, q["rpay"].ptr
            a.skey, a.spay = q["skey"].ptr, q["spay"].ptr
            a.sfilt = q["sfilt"].ptr if q.get("sfilt") is not None else None
            a.nr, a.ns = int(q["nr"]), int(q["ns"])
            a.flo, a.fhi = int(q["flo"]), int(q["fhi"])
            a.okey, a.orpay, a.ospay = q["okey"].ptr, q["orpay"].ptr, q["ospay"].ptr
            a.out_capacity = int(q["cap"])
        h = C.c_void_p()
        _lib.call("qo_batch_create", arr, len(queries), C.byref(h))
        self.handle = h.value
        self.nq = len(queries)

    def launch(self, stream=None) -> None:
        _lib.call("qo_batch_launch", self.handle, stream.handle if stream is not None else None)

    def counts(self) -> np.ndarray:
        out = np.zeros(self.nq, np.uint64)
        _lib.call("qo_batch_counts", self.handle, out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def destroy(self) -> None:
        if self.handle:
            _lib.call("qo_batch_destroy", self.handle)
            self.handle = None


def batch_hj_filter_sort(queries, capacities=None):
    """Host-array convenience wrapper: queries are tuples (rkey, rpay, skey, spay,
    sfilt, flo, fhi); returns a list of (key, r_payload, s_payload).  Each query's
    output holds |S| rows unless `capacities` gives its row capacity (duplicate
    R keys can give more matches than S has tuples)."""
    dev, bufs = [], []
    for i, (rk, rp, sk, sp, sf, flo, fhi) in enumerate(queries):
        rk, rp, sk, sp, sf = (_i32(a) for a in (rk, rp, sk, sp, sf))
        cap = sk.shape[0] if capacities is None else int(capacities[i])
        d = {"rkey": DeviceBuffer.from_numpy(rk), "rpay": DeviceBuffer.from_numpy(rp),
             "skey": DeviceBuffer.from_numpy(sk), "spay": DeviceBuffer.from_numpy(sp),
             "sfilt": DeviceBuffer.from_numpy(sf), "nr": rk.shape[0], "ns": sk.shape[0],
             "flo": flo, "fhi": fhi, "cap": cap}
        for c in ("okey", "orpay", "ospay"):
            d[c] = DeviceBuffer(max(cap, 1) * 4)
        dev.append(d)
        bufs += [v for v in d.values() if isinstance(v, DeviceBuffer)]
    b = JoinBatch(dev)
    try:
        b.launch()
        counts = b.counts()
        for d, m in zip(dev, counts):
            if int(m) > d["cap"]:
                raise _lib.QoError(3, "batch_hj_filter_sort", f"{int(m)} matches exceed capacity {d['cap']}")
        out = [tuple(d[c].to_numpy(np.int32, int(m)) for c in ("okey", "orpay", "ospay"))
               for d, m in zip(dev, counts)]
    finally:
        b.destroy()
        for x in bufs:
            x.free()
    return out


def projection(h_Rin, proj) -> np.ndarray:
    """out[i] = (proj[i].x, Rin[proj[i].x].y) (validateProjection, Validate.cpp:39-71)."""
    r, pj = _records(h_Rin), _records(proj)
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer.from_numpy(pj) as d_p, \
            DeviceBuffer(max(pj.nbytes, 8)) as d_o:
        _lib.call("qo_project_records", d_r.ptr, r.shape[0], d_p.ptr, pj.shape[0], d_o.ptr, None)
        return d_o.to_numpy(np.uint32, pj.size).reshape(-1, 2)


def reduce(h_Rin, op: int) -> int:
    """REDUCE_SUM / MAX / MIN / AVERAGE over y as validateReduce computes it (uint32)."""
    r = _records(h_Rin)
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer(4) as d_out:
        _lib.call("qo_reduce_records", d_r.ptr, r.shape[0], int(op), d_out.ptr, None)
        return int(d_out.to_numpy(np.uint32, 1)[0])


def unpack_records(h_Rin):
    """(x, y) columns of the records."""
    r = _records(h_Rin)
    n = r.shape[0]
    with DeviceBuffer.from_numpy(r) as d_r, DeviceBuffer(max(n, 1) * 4) as dx, DeviceBuffer(max(n, 1) * 4) as dy:
        _lib.call("qo_unpack_records", d_r.ptr, n, dx.ptr, dy.ptr, None)
        return dx.to_numpy(np.uint32, n), dy.to_numpy(np.uint32, n)


def sort_records(h_Rin) -> np.ndarray:
    rin = _records(h_Rin)
    with DeviceBuffer.from_numpy(rin) as d:
        _lib.call("qo_sort_records", d.ptr, rin.shape[0], None)
        return d.to_numpy(np.uint32, rin.size).reshape(-1, 2)


def sort_join_rows(key, rpay, spay):
    k, r, s = _i32(key), _i32(rpay), _i32(spay)
    with DeviceBuffer.from_numpy(k) as dk, DeviceBuffer.from_numpy(r) as dr, DeviceBuffer.from_numpy(s) as ds:
        _lib.call("qo_sort_join_rows", dk.ptr, dr.ptr, ds.ptr, k.shape[0], None)
        n = k.shape[0]
        return dk.to_numpy(np.int32, n), dr.to_numpy(np.int32, n), ds.to_numpy(np.int32, n)


# ---------------------------------------------------------------------------
def _clear_error() -> None:
    # void entry points report failures only through qo_last_error().
    _lib.load().qo_clear_error()


def _raise_if_error(where: str) -> None:
    msg = _lib.last_error()
    if msg:
        raise _lib.QoError(-1, where, msg)
