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
