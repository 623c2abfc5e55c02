"""Device memory, streams and events over libqo's plumbing entry points."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceBuffer:
    """An owned HIP device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _lib.call("qo_malloc", C.byref(p), max(self.nbytes, 1))
        self._ptr = p.value

    @property
    def ptr(self) -> int:
        if self._ptr is None:
            raise ValueError("buffer already freed")
        return self._ptr

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            _lib.call("qo_memcpy_h2d", buf.ptr, arr.ctypes.data, arr.nbytes)
        return buf

    def to_numpy(self, dtype, count: int | None = None, shape=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            if out.nbytes > self.nbytes:
                raise ValueError("read beyond the buffer")
            _lib.call("qo_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes)
        return out.reshape(shape) if shape is not None else out

    def write(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("write beyond the buffer")
        if arr.nbytes:
            _lib.call("qo_memcpy_h2d", self.ptr, arr.ctypes.data, arr.nbytes)

    def free(self) -> None:
        if self._ptr is not None:
            _lib.call("qo_free", self._ptr)
            self._ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream created by libqo."""

    def __init__(self):
        h = C.c_void_p()
        _lib.call("qo_stream_create", C.byref(h))
        self.handle = h.value

    def sync(self) -> None:
        _lib.call("qo_stream_sync", self.handle)

    def destroy(self) -> None:
        if self.handle is not None:
            _lib.call("qo_stream_destroy", self.handle)
            self.handle = None


class Event:
    def __init__(self):
        h = C.c_void_p()
        _lib.call("qo_event_create", C.byref(h))
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> None:
        _lib.call("qo_event_record", self.handle, stream.handle if stream else None)

    def sync(self) -> None:
        _lib.call("qo_event_sync", self.handle)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _lib.call("qo_event_elapsed_ms", self.handle, end.handle, C.byref(ms))
        return float(ms.value)

    def destroy(self) -> None:
        if self.handle is not None:
            _lib.call("qo_event_destroy", self.handle)
            self.handle = None


def device_info() -> dict:
    info = _lib.QoDeviceInfo()
    _lib.call("qo_get_device_info", C.byref(info))
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "cu_count": info.cu_count,
        "lds_bytes_per_block": info.lds_bytes_per_block,
        "wave_size": info.wave_size,
        "hbm_bytes": info.hbm_bytes,
        "xcd_count": info.xcd_count,
    }


def synchronize() -> None:
    _lib.call("qo_device_sync")
