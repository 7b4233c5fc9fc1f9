"""Loader of libswift_amd.so.  Fails loudly when the library is missing:
there is no CPU fallback anywhere in the product."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswift_amd.so")

_lib = None


class SwiftError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SwiftError(f"{LIB_PATH} is not built; run `python -m swift_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.swift_last_error.restype = C.c_char_p
        L.swift_version.restype = C.c_char_p
        for name in ("swift_device_alloc", "swift_device_free", "swift_memcpy_htod",
                     "swift_memcpy_dtoh", "swift_memset_device", "swift_device_synchronize",
                     "swift_device_count", "swift_sw_batch", "swift_align_sequences",
                     "swift_sw_batch_stats", "swift_set_tuning", "swift_set_coarse_rows",
                     "swift_get_coarse_rows", "swift_set_map_mode", "swift_get_map_mode",
                     "swift_set_plain_chunk"):
            getattr(L, name).restype = C.c_int
        L.swift_device_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.swift_device_free.argtypes = [C.c_void_p]
        L.swift_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.swift_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.swift_memset_device.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise SwiftError(lib().swift_last_error().decode(errors="replace"))


def exported_symbols():
    """Names declared in include/swift_abi.h (for the export test)."""
    import re
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "swift_abi.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", hdr)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.isupper()})
