"""MI355X (gfx950) relational operator library: filter, group-by aggregation and
hash join + sort, as hand-written HIP kernels behind a C ABI (libqo.so).

Python is a thin host: ``operators`` mirrors the reference's operator entry
points, ``device`` wraps device memory/streams/events, and ``data`` builds the
benchmark tables on the device.  Importing the package loads libqo.so and fails
if it has not been built.
"""
from __future__ import annotations

from . import _lib

_lib.load()

from . import device, operators  # noqa: E402

__all__ = ["device", "operators", "lib"]


def lib():
    """The loaded libqo ctypes handle."""
    return _lib.load()
