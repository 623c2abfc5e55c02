"""Synthetic benchmark tables generated directly in device memory.

Every column is a deterministic function of (seed, stream, row) (see
csrc/qo_rng.hpp), so any prefix of a table can be regenerated bit-identically
on the host for checking.  Seeds follow ``SEED_BASE + config number``.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceBuffer

SEED_BASE = 0x51A7E5EED0000000
INT_MIN, INT_MAX = -(1 << 31), (1 << 31) - 1

# Column stream ids inside one configuration's seed.
ST_RKEY, ST_RPAY, ST_SKEY, ST_SPAY, ST_SFILT, ST_KEY, ST_VAL = 0, 1, 2, 3, 4, 5, 6


def gen_uniform(n: int, seed: int, stream: int, lo: int, rng: int, s=None) -> DeviceBuffer:
    b = DeviceBuffer(max(n, 1) * 4)
    _lib.call("qo_gen_uniform_u32", b.ptr, n, seed, stream, lo, rng, s)
    return b


def gen_perm(n: int, seed: int, s=None) -> DeviceBuffer:
    b = DeviceBuffer(max(n, 1) * 4)
    _lib.call("qo_gen_perm_u32", b.ptr, n, seed, s)
    return b


def gen_fk(n: int, seed: int, stream: int, nr: int, bits: int, key_seed: int, s=None) -> DeviceBuffer:
    b = DeviceBuffer(max(n, 1) * 4)
    _lib.call("qo_gen_fk_u32", b.ptr, n, seed, stream, nr, bits, key_seed, s)
    return b


def gen_bij(n: int, bits: int, seed: int, s=None) -> DeviceBuffer:
    b = DeviceBuffer(max(n, 1) * 4)
    _lib.call("qo_gen_bij_u32", b.ptr, n, bits, seed, s)
    return b


def gen_zipf(n: int, G: int, z: float, seed: int, stream: int, perm_seed: int, s=None) -> DeviceBuffer:
    t = np.empty(G, np.uint64)
    _lib.call("qo_zipf_table", t.ctypes.data, G, float(z))
    b = DeviceBuffer(max(n, 1) * 4)
    with DeviceBuffer.from_numpy(t) as dt:
        _lib.call("qo_gen_zipf_u32", b.ptr, n, dt.ptr, G, seed, stream, perm_seed, s)
        _lib.call("qo_device_sync")
    return b


class JoinTables:
    """Filtered hash-join inputs: R (key = seeded permutation of [0, nr), payload
    uniform 32-bit) and S (fk uniform over R's keys, payload uniform 32-bit,
    filter column uniform in [0, 100))."""

    def __init__(self, nr: int, ns: int, seed: int, filt_range: int = 100):
        self.nr, self.ns, self.seed = nr, ns, seed
        self.rkey = gen_perm(nr, seed + 1)
        self.rpay = gen_uniform(nr, seed, ST_RPAY, 0, 0)
        self.skey = gen_fk(ns, seed, ST_SKEY, nr, 0, seed + 1)
        self.spay = gen_uniform(ns, seed, ST_SPAY, 0, 0)
        self.sfilt = gen_uniform(ns, seed, ST_SFILT, 0, filt_range)
        _lib.call("qo_device_sync")

    def free(self):
        for b in (self.rkey, self.rpay, self.skey, self.spay, self.sfilt):
            b.free()


class FilterAggTables:
    """key uniform in [0, 2^31), value uniform in [0, 1000)."""

    def __init__(self, n: int, seed: int):
        self.n, self.seed = n, seed
        self.key = gen_uniform(n, seed, ST_KEY, 0, 1 << 31)
        self.val = gen_uniform(n, seed, ST_VAL, 0, 1000)
        _lib.call("qo_device_sync")

    def free(self):
        self.key.free()
        self.val.free()


class GroupbyTables:
    """Group key Zipf(z) over G ranks mapped through a seeded permutation of
    [0, G); value uniform in [0, 2^16)."""

    def __init__(self, n: int, G: int, z: float, seed: int):
        self.n, self.G, self.z, self.seed = n, G, z, seed
        self.key = gen_zipf(n, G, z, seed, ST_KEY, seed + 2)
        self.val = gen_uniform(n, seed, ST_VAL, 0, 1 << 16)
        _lib.call("qo_device_sync")

    def free(self):
        self.key.free()
        self.val.free()
