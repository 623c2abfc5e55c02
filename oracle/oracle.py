"""ctypes wrapper of the CPU oracle (oracle/qo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None
vp = C.c_void_p
sz = C.c_size_t
u64, i64, u32, i32 = C.c_uint64, C.c_int64, C.c_uint32, C.c_int32

_PROTOS = {
    "or_mix64": (u64, [u64]),
    "or_rand64": (u64, [u64, u64, u64]),
    "or_gen_uniform_u32": (None, [vp, sz, u64, u64, u32, u32]),
    "or_bij": (u32, [u32, u32, u64]),
    "or_perm": (u32, [u64, u64, u64]),
    "or_gen_perm": (None, [vp, sz, u64]),
    "or_zipf_table": (None, [vp, u32, C.c_double]),
    "or_gen_zipf_keys": (None, [vp, sz, vp, u32, u64, u64, u64]),
    "or_scan_excl_i32": (None, [vp, vp, sz]),
    "or_filter_records": (u64, [vp, sz, i32, i32, vp]),
    "or_filter_agg_records": (None, [vp, sz, i32, i32, vp]),
    "or_filter_agg": (None, [vp, vp, sz, i32, i32, vp]),
    "or_sort_records": (None, [vp, sz]),
    "or_group_startpos": (u32, [vp, sz, vp]),
    "or_parallel_aggregate": (None, [vp, sz, vp, u32, C.c_int, vp]),
    "or_groupby_agg": (sz, [vp, vp, sz, vp, vp, vp, vp, vp]),
    "or_hash_join": (sz, [vp, vp, sz, vp, vp, vp, sz, i32, i32, vp, vp, vp, sz]),
    "or_sort_join_rows": (None, [vp, vp, vp, sz]),
    "or_probe_pairs": (sz, [vp, sz, vp, sz, vp, sz]),
    "or_split_records": (None, [vp, sz, u32, vp, vp]),
    "or_gather_records": (None, [vp, sz, vp, sz, vp]),
    "or_scatter_records": (None, [vp, sz, vp, vp]),
    "or_project_records": (None, [vp, sz, vp, sz, vp]),
    "or_reduce_records": (u32, [vp, sz, C.c_int]),
    "or_unpack_records": (None, [vp, sz, vp, vp]),
    # multi-threaded restatement (qo_cpu_par.c): the timed CPU baseline
    "pr_threads": (C.c_int, [C.c_int]),
    "pr_filter_agg": (None, [vp, vp, sz, i32, i32, vp, C.c_int]),
    "pr_groupby_agg": (sz, [vp, vp, sz, vp, vp, vp, vp, vp, C.c_int]),
    "pr_hash_join": (sz, [vp, vp, sz, vp, vp, vp, sz, i32, i32, vp, vp, vp, sz, C.c_int]),
    "pr_sort_join_rows": (None, [vp, vp, vp, sz, C.c_int]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        for name, (res, args) in _PROTOS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---- generator ---------------------------------------------------------------
def rand64(seed: int, stream: int, i: int) -> int:
    return int(load().or_rand64(seed, stream, i))


def gen_uniform_u32(n: int, seed: int, stream: int, lo: int, rng: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    load().or_gen_uniform_u32(_p(out), n, seed, stream, lo, rng)
    return out


def bij(x: int, bits: int, seed: int) -> int:
    return int(load().or_bij(x, bits, seed))


def perm(i: int, n: int, seed: int) -> int:
    return int(load().or_perm(i, n, seed))


def gen_perm(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    load().or_gen_perm(_p(out), n, seed)
    return out


def zipf_table(G: int, z: float) -> np.ndarray:
    t = np.empty(G, dtype=np.uint64)
    load().or_zipf_table(_p(t), G, z)
    return t


def gen_zipf_keys(n: int, table: np.ndarray, seed: int, stream: int, perm_seed: int) -> np.ndarray:
    t = _c(table, np.uint64)
    out = np.empty(n, dtype=np.uint32)
    load().or_gen_zipf_keys(_p(out), n, _p(t), t.shape[0], seed, stream, perm_seed)
    return out


# ---- operators ---------------------------------------------------------------
def scan_excl_i32(x) -> np.ndarray:
    a = _c(x, np.int32)
    out = np.empty_like(a)
    load().or_scan_excl_i32(_p(a), _p(out), a.shape[0])
    return out


def filter_records(rec, lo: int, hi: int) -> np.ndarray:
    r = _c(rec, np.uint32).reshape(-1, 2)
    out = np.empty_like(r)
    m = load().or_filter_records(_p(r), r.shape[0], lo, hi, _p(out))
    return out[:m].copy()


def filter_agg(key, val, lo: int, hi: int):
    k, v = _c(key, np.int32), _c(val, np.int32)
    out = np.zeros(3, dtype=np.int64)
    load().or_filter_agg(_p(k), _p(v), k.shape[0], lo, hi, _p(out))
    return int(out[0]), int(out[1]), int(out[2])


def filter_agg_records(rec, lo: int, hi: int):
    r = _c(rec, np.uint32).reshape(-1, 2)
    out = np.zeros(3, dtype=np.int64)
    load().or_filter_agg_records(_p(r), r.shape[0], lo, hi, _p(out))
    return int(out[0]), int(out[1]), int(out[2])


def sort_records(rec) -> np.ndarray:
    r = _c(rec, np.uint32).reshape(-1, 2).copy()
    load().or_sort_records(_p(r), r.shape[0])
    return r


def group_startpos(sorted_rec) -> np.ndarray:
    r = _c(sorted_rec, np.uint32).reshape(-1, 2)
    sp = np.empty(max(r.shape[0], 1), dtype=np.int32)
    g = load().or_group_startpos(_p(r), r.shape[0], _p(sp))
    return sp[:g].copy()


def parallel_aggregate(sorted_rec, start_pos, op: int) -> np.ndarray:
    r = _c(sorted_rec, np.uint32).reshape(-1, 2)
    sp = _c(start_pos, np.int32)
    out = np.empty(sp.shape[0], dtype=np.int32)
    load().or_parallel_aggregate(_p(r), r.shape[0], _p(sp), sp.shape[0], op, _p(out))
    return out


def groupby_agg(key, val):
    k, v = _c(key, np.int32), _c(val, np.int32)
    n = k.shape[0]
    ok = np.empty(max(n, 1), np.int32)
    oc = np.empty(max(n, 1), np.int64)
    osum = np.empty(max(n, 1), np.int64)
    omn = np.empty(max(n, 1), np.int32)
    omx = np.empty(max(n, 1), np.int32)
    g = load().or_groupby_agg(_p(k), _p(v), n, _p(ok), _p(oc), _p(osum), _p(omn), _p(omx))
    return ok[:g].copy(), oc[:g].copy(), osum[:g].copy(), omn[:g].copy(), omx[:g].copy()


def hash_join(rkey, rpay, skey, spay, sfilt=None, flo: int = -(1 << 31), fhi: int = (1 << 31) - 1):
    rk, rp, sk, sp = (_c(a, np.int32) for a in (rkey, rpay, skey, spay))
    sf = None if sfilt is None else _c(sfilt, np.int32)
    lib = load()
    m = lib.or_hash_join(_p(rk), _p(rp), rk.shape[0], _p(sk), _p(sp), _p(sf), sk.shape[0], flo, fhi,
                         None, None, None, 0)
    ok, orr, os_ = (np.empty(max(m, 1), np.int32) for _ in range(3))
    lib.or_hash_join(_p(rk), _p(rp), rk.shape[0], _p(sk), _p(sp), _p(sf), sk.shape[0], flo, fhi,
                     _p(ok), _p(orr), _p(os_), m)
    return ok[:m].copy(), orr[:m].copy(), os_[:m].copy()


def sort_join_rows(key, rpay, spay):
    k, r, s = (_c(a, np.int32).copy() for a in (key, rpay, spay))
    load().or_sort_join_rows(_p(k), _p(r), _p(s), k.shape[0])
    return k, r, s


def probe_pairs(r_table, s_table) -> np.ndarray:
    r = _c(r_table, np.uint32).reshape(-1)
    s = _c(s_table, np.uint32).reshape(-1)
    lib = load()
    m = lib.or_probe_pairs(_p(r), r.shape[0] // 2, _p(s), s.shape[0] // 2, None, 0)
    out = np.empty(max(2 * m, 2), np.uint32)
    lib.or_probe_pairs(_p(r), r.shape[0] // 2, _p(s), s.shape[0] // 2, _p(out), m)
    return out[: 2 * m].reshape(m, 2).copy()


# ---- split / scatter / gather / projection / reduce ---------------------------
def split_records(rec, num_part: int):
    """(out, bound): stable split by .y % num_part; bound[p] = (start, end)."""
    r = _c(rec, np.uint32).reshape(-1, 2)
    out = np.empty_like(r)
    bound = np.empty((num_part, 2), np.uint32)
    load().or_split_records(_p(r), r.shape[0], num_part, _p(out), _p(bound))
    return out, bound


def gather_records(rec, loc, s_init=None) -> np.ndarray:
    r = _c(rec, np.uint32).reshape(-1, 2)
    lc = _c(loc, np.int32)
    s = np.zeros((lc.shape[0], 2), np.uint32) if s_init is None else _c(s_init, np.uint32).reshape(-1, 2).copy()
    load().or_gather_records(_p(r), r.shape[0], _p(lc), lc.shape[0], _p(s))
    return s


def scatter_records(rec, loc, s_init=None) -> np.ndarray:
    r = _c(rec, np.uint32).reshape(-1, 2)
    lc = _c(loc, np.int32)
    s = np.zeros_like(r) if s_init is None else _c(s_init, np.uint32).reshape(-1, 2).copy()
    load().or_scatter_records(_p(r), r.shape[0], _p(lc), _p(s))
    return s


def project_records(rec, proj) -> np.ndarray:
    r = _c(rec, np.uint32).reshape(-1, 2)
    pj = _c(proj, np.uint32).reshape(-1, 2)
    out = np.empty_like(pj)
    load().or_project_records(_p(r), r.shape[0], _p(pj), pj.shape[0], _p(out))
    return out


def reduce_records(rec, op: int) -> int:
    r = _c(rec, np.uint32).reshape(-1, 2)
    return int(load().or_reduce_records(_p(r), r.shape[0], op))


def unpack_records(rec):
    r = _c(rec, np.uint32).reshape(-1, 2)
    x = np.empty(r.shape[0], np.uint32)
    y = np.empty(r.shape[0], np.uint32)
    load().or_unpack_records(_p(r), r.shape[0], _p(x), _p(y))
    return x, y
