"""The C-ABI library loads and exports every symbol declared in include/ (CPU)."""
import ctypes
import os

from tdscha_amd import _lib


def test_library_exports_declared_symbols():
    assert os.path.exists(_lib.LIB_PATH), "libtdscha_amd.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.declared_symbols()
    assert len(names) >= 6
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_reporting_without_device():
    L = _lib.lib()
    rc = L.tdscha_ctx_create(None, 0, 0, None, None, None, None, 0.0, 0, 0, None, None, None,
                             None, 0)
    assert rc == 2
    assert b"NULL" in L.tdscha_last_error() or b"positive" in L.tdscha_last_error()
