"""roctx ranges around the public entry points (the reference's NVTX ranges
of its I/O library, cpp/include/cucim/profiler/nvtx3.h, for rocprofv3's
``--marker-trace``).

Off unless ``CUCIM_AMD_ROCTX=1`` is set when the package is imported (or
:func:`enable` is called): then each annotated call pushes a range named
``cucim_amd.<name>`` on the calling thread and pops it on return.  The roctx
library is loaded on first use; without it the ranges are no-ops.
"""

import ctypes
import functools
import os

_enabled = os.environ.get("CUCIM_AMD_ROCTX", "") not in ("", "0")
_lib = None
_tried = False
_NAMES = ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4",
          "libroctx64.so")


def _roctx():
    global _lib, _tried
    if not _tried:
        _tried = True
        for name in _NAMES:
            for cand in (name, os.path.join("/opt/rocm/lib", name)):
                try:
                    lib = ctypes.CDLL(cand)
                except OSError:
                    continue
                lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                lib.roctxRangePushA.restype = ctypes.c_int
                lib.roctxRangePop.argtypes = []
                lib.roctxRangePop.restype = ctypes.c_int
                _lib = lib
                return _lib
    return _lib


def enable(on=True):
    """Switch the ranges on or off; returns whether a roctx library is
    available."""
    global _enabled
    _enabled = bool(on)
    return _roctx() is not None


def enabled():
    return _enabled


class range_:
    """``with range_("name"):`` pushes and pops one roctx range."""

    __slots__ = ("msg", "lib")

    def __init__(self, name):
        self.msg = ("cucim_amd." + name).encode()
        self.lib = None

    def __enter__(self):
        if _enabled:
            self.lib = _roctx()
            if self.lib is not None:
                self.lib.roctxRangePushA(self.msg)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxRangePop()
            self.lib = None
        return False


def annotate(name):
    """Decorator: the call runs inside a roctx range ``cucim_amd.<name>``."""
    msg = ("cucim_amd." + name).encode()

    def deco(fn):
        @functools.wraps(fn)
        def wrapper(*args, **kwargs):
            if not _enabled:
                return fn(*args, **kwargs)
            lib = _roctx()
            if lib is None:
                return fn(*args, **kwargs)
            lib.roctxRangePushA(msg)
            try:
                return fn(*args, **kwargs)
            finally:
                lib.roctxRangePop()

        return wrapper

    return deco
