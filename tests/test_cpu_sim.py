"""CPU tests that run libqo's own kernels on the CPU build (tests/cpu_sim): the
csrc sources compiled for the host with work-items as fibers, so host planning
and kernel logic are exercised without a GPU.

* runtime_checks: per-thread workspaces of the NULL stream; the bounded
  look-back wait raising QO_ERR_DEVICE_WAIT when its predecessor never publishes.
* serve_mix: the serving tests' four tables (scaled to 2%) and 24 mixed requests
  in batches of 8 through the serving loop, every result equal to the same
  request run alone on the plain path; with canaries after every buffer carved
  from a batch's arena, no kernel writes past its buffer.
* record_ops: split (one and two radix passes), gather, scatter, projection,
  reduce and Records -> columns on the benchmark's input shape, each against a
  plain loop.
* range_build: the key-range join table built slice by slice in LDS (16384 full
  slots per slice, or, from 2^21 slots, 2^15 payloads and presence bits per
  slice after one split by key) equals the table built by storing every tuple
  at its slot (permutation keys, a negative
  base with (-1, -1) tuples, keys with gaps, a duplicate key's hash fallback).
* range_probe: joins whose key-range table is probed through LDS slices (the
  default plan from 2^21 slots), through L2 and unpartitioned, each equal to a
  host hash join (a permutation, sparse keys in a larger table, a (-1, -1) build
  tuple, a table below the LDS route's minimum), and qo_hj_filter_sort's rows in
  order.
* batch_parts: a batch of ten joins (one with an empty build side, one with a
  single row) with the partitioned build and probe of large tables by their
  rule, forced for every query and switched off, each query's rows equal to the
  same query run alone through qo_hj_filter_sort.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "cpu_sim"


@pytest.fixture(scope="module")
def cpu_build():
    env = dict(os.environ, SIMT_VARIANT="fast")
    subprocess.run([sys.executable, str(HERE / "build.py"), "serve_mix", "runtime_checks", "record_ops", "batch_parts", "range_build", "range_probe"],
                   check=True, env=env, timeout=900)
    return HERE / "_build" / "fast"


def _run(exe, *args, timeout=600):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("PASS"), r.stdout[-4000:]
    return r.stdout


def test_runtime_checks(cpu_build):
    _run(cpu_build / "runtime_checks")


def test_served_mix_equals_alone_cpu(cpu_build):
    out = _run(cpu_build / "serve_mix", 0.02, 24)
    assert out.count(" ok") == 24


def test_served_mix_writes_stay_in_their_buffers_cpu(cpu_build):
    """The same mix one batch at a time, with a canary after every buffer carved
    from the batch's arena (result slots, filter descriptors, group-by outputs
    and workspaces, join outputs, the join plan and its sort and partition
    workspaces): no kernel writes past its buffer.  canary-part forces the
    partitioned join build and probe for every table."""
    for mode in ("canary", "canary-part"):
        out = _run(cpu_build / "serve_mix", 0.05, 24, mode)
        assert "canaries: 0 damaged" in out and out.count(" ok") == 24


def test_record_ops_cpu(cpu_build):
    out = _run(cpu_build / "record_ops", 16)
    assert "split 32 partitions: 0 mismatches" in out and "split 1000 partitions: 0 mismatches" in out


def test_batch_parts_cpu(cpu_build):
    out = _run(cpu_build / "batch_parts", 0.25)
    assert out.count("0 of 10 queries differ") == 3


def test_range_build_sliced_equals_direct_cpu(cpu_build):
    out = _run(cpu_build / "range_build", 20)  # 2^20 slots: 16384-slot slices of full slots
    assert out.count("same table") == 4
    out = _run(cpu_build / "range_build", 21)  # 2^21 slots and more: 2^15-slot slices built in LDS
    assert out.count("same table") == 4


def test_range_probe_routes_equal_host_join_cpu(cpu_build):
    out = _run(cpu_build / "range_probe", 20)
    assert out.count("default plan same rows, L2 same rows, direct same rows") == 4
    assert "hj_filter_sort same rows in order" in out
