"""Minimal device-memory plumbing over the library's allocation helpers
(no PyTorch needed on the hot path)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceBuffer:
    """A HIP device allocation of `nbytes`."""

    def __init__(self, nbytes: int):
        self.nbytes = int(max(nbytes, 1))
        p = C.c_void_p()
        check(lib().swift_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().swift_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(lib().swift_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().swift_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self) -> None:
        check(lib().swift_memset_device(self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr:
            check(lib().swift_device_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def cptr(self):
        return C.c_void_p(self.ptr)


class Stream:
    """A non-blocking HIP stream of the library (for work that overlaps other
    streams' work, such as pipeline graphs launched in turn)."""

    def __init__(self):
        p = C.c_void_p()
        check(lib().swift_stream_create(C.byref(p)))
        self.ptr = p.value

    def synchronize(self) -> None:
        check(lib().swift_stream_synchronize(C.c_void_p(self.ptr)))

    def close(self) -> None:
        if self.ptr:
            check(lib().swift_stream_destroy(C.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cptr(self):
        return C.c_void_p(self.ptr)


def synchronize() -> None:
    check(lib().swift_device_synchronize())


def device_count() -> int:
    n = C.c_int()
    check(lib().swift_device_count(C.byref(n)))
    return n.value
