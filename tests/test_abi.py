"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/swift_abi.h declares."""
import ctypes as C

import pytest


@pytest.fixture(scope="module")
def built():
    from swift_amd import _lib
    import os
    if not os.path.exists(_lib.LIB_PATH):
        from swift_amd import build
        build.build()
    return _lib


def test_library_exports_header_symbols(built):
    L = built.lib()
    names = built.exported_symbols()
    assert "alignSequences" in names and "swift_sw_batch" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_without_gpu(built):
    assert built.lib().swift_version().startswith(b"swift_amd")


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from swift_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.SwiftError):
        _lib.lib()


def test_mode_switches_validate_without_gpu(built):
    """The process-wide switches (row classes, mapping entry tier, tuning)
    take their documented ranges, reject others with a message, and touch no
    device."""
    L = built.lib()
    m = C.c_int()
    assert L.swift_get_map_mode(C.byref(m)) == 0
    before_map = m.value
    assert L.swift_get_coarse_rows(C.byref(m)) == 0
    before_rows = m.value
    w, i16, wd = C.c_int(), C.c_int(), C.c_int()
    assert L.swift_get_tuning(C.byref(w), C.byref(i16), C.byref(wd)) == 0
    before_tuning = (w.value, i16.value, wd.value)
    try:
        for mode in (-1, 0, 1):
            assert L.swift_set_map_mode(mode) == 0
            assert L.swift_get_map_mode(C.byref(m)) == 0 and m.value == mode
            assert L.swift_set_coarse_rows(mode) == 0
            assert L.swift_get_coarse_rows(C.byref(m)) == 0 and m.value == mode
        for bad in (-2, 2, 99):
            assert L.swift_set_map_mode(bad) != 0
            assert b"map mode" in L.swift_last_error()
            assert L.swift_set_coarse_rows(bad) != 0
            assert b"row class" in L.swift_last_error()
        assert L.swift_set_plain_chunk(-1) != 0
        assert b"chunk" in L.swift_last_error()
        assert L.swift_set_plain_chunk(7) == 0 and L.swift_set_plain_chunk(0) == 0
        assert L.swift_set_tuning(65, 1, 1) != 0
        assert L.swift_set_tuning(-1, 1, 1) != 0
        assert L.swift_set_tuning(8, 1, 0) == 0
        assert L.swift_get_tuning(C.byref(w), C.byref(i16), C.byref(m)) == 0
        assert (w.value, i16.value, m.value) == (8, 1, 0)
    finally:
        L.swift_set_map_mode(before_map)
        L.swift_set_coarse_rows(before_rows)
        L.swift_set_tuning(*before_tuning)
