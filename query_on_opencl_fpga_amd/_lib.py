"""ctypes bindings of libqo.so (the C ABI declared in include/qo.h and include/qo_compat.h).

The library is built in-tree by ``query_on_opencl_fpga_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any operator.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libqo.so"

c_size = C.c_size_t
c_u64 = C.c_uint64
c_i64 = C.c_int64
c_u32 = C.c_uint32
c_i32 = C.c_int32
vp = C.c_void_p
P = C.POINTER


class QoRecord(C.Structure):
    _fields_ = [("x", c_u32), ("y", c_u32)]


class QoDeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("arch", C.c_char * 32),
        ("cu_count", C.c_int),
        ("lds_bytes_per_block", C.c_int),
        ("wave_size", C.c_int),
        ("hbm_bytes", c_size),
        ("xcd_count", C.c_int),
    ]


class QoJoinQuery(C.Structure):
    _fields_ = [
        ("rkey", vp), ("rpay", vp),
        ("skey", vp), ("spay", vp), ("sfilt", vp),
        ("nr", c_u64), ("ns", c_u64),
        ("flo", c_i32), ("fhi", c_i32),
        ("okey", vp), ("orpay", vp), ("ospay", vp),
        ("out_capacity", c_u64),
    ]


class ScanPara(C.Structure):
    # host/host/testScan.h:4-12 field order, then the single-pass scan's workspace
    _fields_ = [("outputBuffer", vp), ("blockSumBuffer", vp), ("tempBuffer", vp), ("pass_", C.c_uint),
                ("gLength", C.c_uint), ("blockSize", C.c_uint), ("ws", vp)]


class QoQuery(C.Structure):
    _fields_ = [("filter", C.c_int), ("lo", c_i32), ("hi", c_i32), ("group", C.c_int), ("op", C.c_int),
                ("selectivity", C.c_double), ("groups", C.c_double)]


PLAN_MAX_SEGS = 4


class QoPlan(C.Structure):
    _fields_ = [("nseg", C.c_int), ("impl", C.c_int * PLAN_MAX_SEGS), ("first_step", C.c_int * PLAN_MAX_SEGS),
                ("last_step", C.c_int * PLAN_MAX_SEGS), ("seg_ms", C.c_double * PLAN_MAX_SEGS),
                ("est_ms", C.c_double), ("tune_op", C.c_int * PLAN_MAX_SEGS),
                ("block", C.c_int * PLAN_MAX_SEGS), ("items", C.c_int * PLAN_MAX_SEGS),
                ("param", C.c_int * PLAN_MAX_SEGS)]


# name -> (restype, argtypes).  restype C.c_int means a qo_status.
PROTOS = {
    # runtime
    "qo_version": (C.c_char_p, []),
    "qo_last_error": (C.c_char_p, []),
    "qo_clear_error": (None, []),
    "qo_init": (C.c_int, [C.c_int]),
    "qo_get_device_info": (C.c_int, [P(QoDeviceInfo)]),
    "qo_malloc": (C.c_int, [P(vp), c_size]),
    "qo_free": (C.c_int, [vp]),
    "qo_memcpy_h2d": (C.c_int, [vp, vp, c_size]),
    "qo_memcpy_d2h": (C.c_int, [vp, vp, c_size]),
    "qo_memcpy_d2d": (C.c_int, [vp, vp, c_size, vp]),
    "qo_fill_u32": (C.c_int, [vp, c_u32, c_size, vp]),
    "qo_copy": (C.c_int, [vp, vp, c_size, vp]),
    "qo_copy_nt": (C.c_int, [vp, vp, c_size, vp]),
    "qo_stream_create": (C.c_int, [P(vp)]),
    "qo_stream_destroy": (C.c_int, [vp]),
    "qo_stream_sync": (C.c_int, [vp]),
    "qo_device_sync": (C.c_int, []),
    "qo_event_create": (C.c_int, [P(vp)]),
    "qo_event_destroy": (C.c_int, [vp]),
    "qo_event_record": (C.c_int, [vp, vp]),
    "qo_event_sync": (C.c_int, [vp]),
    "qo_event_query": (C.c_int, [vp]),
    "qo_event_elapsed_ms": (C.c_int, [vp, vp, P(C.c_float)]),
    "qo_release_workspaces": (C.c_int, []),
    "qo_debug_inject_fault": (C.c_int, [C.c_int]),
    "qo_debug_set_spin_limit": (C.c_int, [c_u32]),
    # generators
    "qo_gen_uniform_u32": (C.c_int, [vp, c_size, c_u64, c_u64, c_u32, c_u32, vp]),
    "qo_gen_perm_u32": (C.c_int, [vp, c_size, c_u64, vp]),
    "qo_gen_bij_u32": (C.c_int, [vp, c_size, c_u32, c_u64, vp]),
    "qo_gen_fk_u32": (C.c_int, [vp, c_size, c_u64, c_u64, c_u64, c_u32, c_u64, vp]),
    "qo_zipf_table": (C.c_int, [vp, c_u32, C.c_double]),
    "qo_gen_zipf_u32": (C.c_int, [vp, c_size, vp, c_u32, c_u64, c_u64, c_u64, vp]),
    "qo_pack_records": (C.c_int, [vp, vp, c_size, vp, vp]),
    "qo_unpack_records": (C.c_int, [vp, c_size, vp, vp, vp]),
    "qo_split_records": (C.c_int, [vp, c_size, c_u32, vp, vp, vp]),
    "qo_gather_records": (C.c_int, [vp, c_size, vp, c_size, vp, vp]),
    "qo_scatter_records": (C.c_int, [vp, c_size, vp, vp, vp]),
    "qo_project_records": (C.c_int, [vp, c_size, vp, c_size, vp, vp]),
    "qo_reduce_records": (C.c_int, [vp, c_size, C.c_int, vp, vp]),
    "qo_plan_query": (C.c_int, [P(QoQuery), c_size, C.c_uint, P(QoPlan)]),
    "qo_plan_cost_set": (C.c_int, [C.c_int, C.c_double, C.c_double]),
    "qo_plan_cost_get": (C.c_int, [C.c_int, P(C.c_double), P(C.c_double)]),
    "qo_plan_calibrate": (C.c_int, [C.c_int, C.c_int]),
    "qo_plan_execute": (C.c_int, [P(QoPlan), P(QoQuery), vp, c_size, c_size, vp, vp, vp, vp, vp]),
    # filter
    "qo_filter_agg": (C.c_int, [vp, vp, c_size, c_i32, c_i32, vp, vp]),
    "qo_filter_agg_records": (C.c_int, [vp, c_size, c_i32, c_i32, vp, vp]),
    "qo_filter_records": (C.c_int, [vp, c_size, c_i32, c_i32, vp, vp, vp]),
    "qo_filter_records_noalloc": (C.c_int, [vp, c_size, c_i32, c_i32, vp, vp, vp]),
    "qo_scan_exclusive_i32": (C.c_int, [vp, vp, c_size, vp]),
    # sort
    "qo_sort_records": (C.c_int, [vp, c_size, vp]),
    "qo_sort_join_rows": (C.c_int, [vp, vp, vp, c_size, vp]),
    # sort-based group-by
    "qo_group_startpos": (C.c_int, [vp, c_size, vp, vp, vp]),
    "qo_aggregate_groups": (C.c_int, [vp, c_size, vp, c_size, C.c_int, vp, vp]),
    # hash group-by
    "qo_groupby_agg": (C.c_int, [vp, vp, c_size, c_size, vp, vp, vp, vp, vp, vp, vp]),
    "qo_groupby_agg_filtered": (C.c_int, [vp, vp, vp, c_size, c_i32, c_i32, c_size, vp, vp, vp, vp, vp, vp, vp]),
    # join
    "qo_join_table_bytes": (c_size, [c_size]),
    "qo_join_strategy": (C.c_int, [c_size, c_size, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
    "qo_join_table_info": (C.c_int, [vp, P(C.c_int), P(C.c_uint64), P(C.c_int), vp]),
    "qo_hash_build": (C.c_int, [vp, c_size, vp, vp, c_size, vp]),
    "qo_hash_probe": (C.c_int, [vp, c_size, vp, vp, vp, c_size, c_i32, c_i32, vp, vp, vp, c_size, vp, vp]),
    "qo_hj_filter_sort": (C.c_int, [vp, c_size, vp, vp, c_size, vp, vp, vp, c_size, c_i32, c_i32,
                                    vp, vp, vp, c_size, P(c_u64), vp]),
    "qo_join_rows_digest": (C.c_int, [vp, vp, vp, c_size, vp, vp, vp]),
    "qo_bucket_build": (C.c_int, [vp, c_u32, vp, c_u32, c_u32, vp, vp]),
    "qo_bucket_probe": (C.c_int, [vp, vp, c_u32, c_u32, vp, c_u32, vp, c_size, vp, vp]),
    # batched
    "qo_batch_create": (C.c_int, [P(QoJoinQuery), C.c_int, P(vp)]),
    "qo_batch_launch": (C.c_int, [vp, vp]),
    "qo_batch_counts": (C.c_int, [vp, P(c_u64)]),
    "qo_batch_destroy": (C.c_int, [vp]),
    # autotuner
    "qo_tune_get": (C.c_int, [C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
    "qo_tune_set": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "qo_tune_load": (C.c_int, [C.c_char_p]),
    "qo_tune_save": (C.c_int, [C.c_char_p]),
    "qo_tune_reset": (C.c_int, []),
    "qo_autotune": (C.c_int, [C.c_int, C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_double),
                              P(C.c_double)]),
    # server prototypes: serving.py
    # compat (reference-named)
    "qo_mem_create": (vp, [c_size]),
    "qo_mem_wrap": (vp, [vp, c_size]),
    "qo_mem_ptr": (vp, [vp]),
    "qo_mem_size": (c_size, [vp]),
    "qo_mem_free": (None, [vp]),
    "qo_mem_write": (C.c_int, [vp, vp, c_size]),
    "qo_mem_read": (C.c_int, [vp, vp, c_size]),
    "initScan": (None, [C.c_int, P(ScanPara)]),
    "closeScan": (None, [P(ScanPara)]),
    "scanImpl": (None, [vp, C.c_int, vp, vp, vp, vp, vp, vp, P(ScanPara), C.c_int]),
    "scanImpl_offset": (None, [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, P(ScanPara), C.c_int]),
    "filterImpl": (None, [vp, C.c_int, C.c_int, P(vp), P(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int,
                          vp, vp, vp, vp, vp, C.c_int]),
    "range_selection": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(vp), C.c_int, C.c_int,
                                  vp, vp, vp, vp, vp, C.c_int]),
    "CL_RangeSelection": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(vp), C.c_int, C.c_int, C.c_int]),
    "point_selection": (C.c_int, [vp, C.c_int, C.c_int, P(vp), C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "CL_PointSelection": (C.c_int, [vp, C.c_int, C.c_int, P(vp), C.c_int, C.c_int, C.c_int]),
    "CL_PointSelectionOnly": (C.c_int, [vp, C.c_int, C.c_int, P(vp), C.c_int, C.c_int, C.c_int]),
    "CL_RangeSelectionOnly": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(vp), C.c_int, C.c_int, C.c_int]),
    "splitImpl": (None, [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "scatterImpl": (None, [vp, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "scatterImpl_forPart": (None, [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "gatherImpl": (None, [vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "radixSortImpl_sm": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "CL_RadixSort": (None, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int]),
    "radixSortImpl": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "CL_RadixSortOnly": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "CL_BitonicSortOnly": (None, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int]),
    "groupByImpl": (C.c_int, [vp, C.c_int, vp, P(vp), C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "CL_GroupBy": (C.c_int, [vp, C.c_int, vp, P(vp), C.c_int, C.c_int, C.c_int]),
    "parallelAggregate_init": (None, [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "aggAfterGroupByImpl": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int,
                                   vp, vp, vp, vp, vp, C.c_int]),
    "agg_max_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "agg_min_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "agg_sum_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "agg_avg_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]),
    "CL_agg_max_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int]),
    "CL_agg_min_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int]),
    "CL_agg_sum_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int]),
    "CL_agg_avg_afterGroupBy": (None, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, C.c_int]),
    "buildHashTable": (C.c_int, [vp, vp, c_u32, c_u32, c_u32, vp]),
    "probeHashTable": (C.c_int, [vp, vp, vp, c_u32, c_u32, c_u32, vp]),
}

_lib = None
MISSING: list[str] = []  # declared entry points the loaded library does not export


class QoError(RuntimeError):
    """A libqo call returned a non-zero qo_status."""

    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


def load() -> C.CDLL:
    """Load libqo.so and attach prototypes (raises if the library is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m query_on_opencl_fpga_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, (res, args) in PROTOS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            MISSING.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().qo_last_error()
    return msg.decode() if msg else ""


def check(status: int, where: str) -> None:
    if status != 0:
        raise QoError(status, where, last_error())


def call(name: str, *args):
    """Call a status-returning entry point and raise QoError on failure."""
    fn = getattr(load(), name)
    rc = fn(*args)
    check(rc, name)
    return rc
