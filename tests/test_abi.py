"""The C-ABI library loads and exports every symbol include/clawgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes

from classic_amd import _lib


def test_library_exports_declared_symbols():
    lib = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert lib.clawgpu_abi_version() == 1


def test_bad_arguments_are_errors_not_crashes():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.clawgpu_create(ctypes.byref(h), 3, 2, 0, 2, 10, 10, 99, None, 0)
    assert rc == _lib.ERR_UNSUPPORTED
    assert b"solver" in lib.clawgpu_last_error()
    rc = lib.clawgpu_create(ctypes.byref(h), 4, 2, 0, 2, 10, 10, 2, None, 0)
    assert rc == _lib.ERR_ARG
