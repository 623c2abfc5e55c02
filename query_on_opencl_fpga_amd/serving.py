"""Serving loop over HBM-resident tables (qo_server_* in include/qo.h).

    sv = Server(budget_bytes=200e9, batch=64)
    t = sv.add_table(key=..., fk=..., grp=..., val=..., filt=..., n=N, group_capacity=G)
    rid = sv.submit(FILTER, t, lo=0, hi=99)              # arrives now
    rid2 = sv.submit_at(GROUPBY, t, arrival_ms=5.0)      # arrives at 5 ms on the server clock
    sv.run()                                             # serves both (continuous batching)
    results = sv.poll(timeout_ms=0)

The loop runs on the calling thread: run(), drain() and poll(timeout_ms != 0)
serve what has arrived, one batch at a time on the server's stream.

Request kinds: FILTER (COUNT, SUM(val) where lo <= filt <= hi), GROUPBY (group
by grp: COUNT/SUM/MIN/MAX per group), JOIN_SORT (build on table_a (key, val),
probe table_b (fk, val) filtered by lo <= filt <= hi, sort by (key, s_payload)).
"""
from __future__ import annotations

import ctypes as C

from . import _lib

FILTER, GROUPBY, JOIN_SORT = 0, 1, 2


class TableDesc(C.Structure):
    _fields_ = [("key", C.c_void_p), ("fk", C.c_void_p), ("grp", C.c_void_p), ("val", C.c_void_p),
                ("filt", C.c_void_p), ("n", C.c_uint64), ("group_capacity", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("table_a", C.c_int), ("table_b", C.c_int), ("lo", C.c_int32),
                ("hi", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("kind", C.c_int), ("count", C.c_uint64),
                ("sum", C.c_int64), ("checksum", C.c_int64), ("submit_ms", C.c_double),
                ("start_ms", C.c_double), ("end_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("max_inflight_batches", C.c_int),
                ("max_batch_queries", C.c_int), ("plain", C.c_int), ("batch_timeout_ms", C.c_double)]


_PROTOS = {
    "qo_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "qo_server_add_table": (C.c_int, [C.c_void_p, C.POINTER(TableDesc), C.POINTER(C.c_int)]),
    "qo_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "qo_server_submit_at": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_double, C.POINTER(C.c_uint64)]),
    "qo_server_run": (C.c_int, [C.c_void_p]),
    "qo_server_poll": (C.c_int, [C.c_void_p, C.POINTER(Result), C.c_int, C.c_int]),
    "qo_server_drain": (C.c_int, [C.c_void_p]),
    "qo_server_destroy": (C.c_int, [C.c_void_p]),
    "qo_server_run_plain": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(Result)]),
    "qo_server_now_ms": (C.c_double, [C.c_void_p]),
}


def _lib_server():
    lib = _lib.load()
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def _ptr(b):
    return None if b is None else (b.ptr if hasattr(b, "ptr") else int(b))


class Server:
    def __init__(self, budget_bytes: float = 200e9, inflight: int = 1, batch: int = 64, plain: bool = False,
                 timeout_ms: float = 60000.0):
        self.lib = _lib_server()
        cfg = ServerConfig(int(budget_bytes), int(inflight), int(batch), 1 if plain else 0, float(timeout_ms))
        h = C.c_void_p()
        _lib.call("qo_server_create", C.byref(cfg), C.byref(h))
        self.handle = h.value

    def add_table(self, n: int, key=None, fk=None, grp=None, val=None, filt=None, group_capacity: int = 0) -> int:
        d = TableDesc(_ptr(key), _ptr(fk), _ptr(grp), _ptr(val), _ptr(filt), int(n), int(group_capacity))
        tid = C.c_int()
        _lib.call("qo_server_add_table", self.handle, C.byref(d), C.byref(tid))
        return tid.value

    def submit(self, kind: int, table_a: int, table_b: int = -1, lo: int = -(1 << 31), hi: int = (1 << 31) - 1) -> int:
        r = Request(kind, table_a, table_b, lo, hi)
        rid = C.c_uint64()
        _lib.call("qo_server_submit", self.handle, C.byref(r), C.byref(rid))
        return rid.value

    def submit_at(self, kind: int, table_a: int, table_b: int = -1, lo: int = -(1 << 31),
                  hi: int = (1 << 31) - 1, arrival_ms: float = 0.0) -> int:
        r = Request(kind, table_a, table_b, lo, hi)
        rid = C.c_uint64()
        _lib.call("qo_server_submit_at", self.handle, C.byref(r), C.c_double(arrival_ms), C.byref(rid))
        return rid.value

    def run(self) -> None:
        _lib.call("qo_server_run", self.handle)

    def poll(self, max_results: int = 4096, timeout_ms: int = -1) -> list[dict]:
        buf = (Result * max_results)()
        n = self.lib.qo_server_poll(self.handle, buf, max_results, timeout_ms)
        if n < 0:
            raise _lib.QoError(-n, "qo_server_poll", _lib.last_error())
        return [buf[i].as_dict() for i in range(n)]

    def drain(self) -> None:
        _lib.call("qo_server_drain", self.handle)

    def run_plain(self, kind: int, table_a: int, table_b: int = -1, lo: int = -(1 << 31),
                  hi: int = (1 << 31) - 1) -> dict:
        r = Request(kind, table_a, table_b, lo, hi)
        out = Result()
        _lib.call("qo_server_run_plain", self.handle, C.byref(r), C.byref(out))
        return out.as_dict()

    def now_ms(self) -> float:
        return self.lib.qo_server_now_ms(self.handle)

    def destroy(self) -> None:
        if self.handle:
            _lib.call("qo_server_destroy", self.handle)
            self.handle = None
