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


def test_user_solver_entry_points_check_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.clawgpu_create_user(ctypes.byref(h), None, 2, 10, 10, None, 0) == _lib.ERR_ARG
    assert b"solver" in lib.clawgpu_last_error()
    assert lib.clawgpu_solver2_load(ctypes.byref(h), None, b"x") == _lib.ERR_ARG
    assert lib.clawgpu_solver2_shape(None, None, None, None) == _lib.ERR_ARG
    assert lib.clawgpu_solver2_destroy(None) == _lib.OK
