"""ctypes binding of libclawgpu.so (the C-ABI declared in include/clawgpu.h).

The library is built in-tree by ``python -m classic_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CLAWGPU_LIB") or os.path.join(HERE, "libclawgpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "clawgpu.h")

OK = 0
ERR_ARG = -1
ERR_UNSUPPORTED = -2
ERR_HIP = -3
ERR_NONFINITE = -4

SOLVER_ACOUSTICS = 1
SOLVER_EULER = 2
SOLVER_VC_ACOUSTICS = 3
SOLVER_ADVECTION = 4
SOLVER_VC_ADVECTION = 5
SOLVER_USER = 100

ARITH_EXACT = 0
ARITH_FAST = 1

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_F = C.POINTER(C.c_float)
_V = C.c_void_p
_d = C.c_double
_i = C.c_int

_SIGS = {
    "clawgpu_abi_version": ([], _i),
    "clawgpu_last_error": ([], C.c_char_p),
    "clawgpu_create": ([C.POINTER(_V), _i, _i, _i, _i, _i, _i, _i, _D, _i], _i),
    "clawgpu_destroy": ([_V], _i),
    "clawgpu_set_q": ([_V, _D], _i),
    "clawgpu_get_q": ([_V, _D], _i),
    "clawgpu_set_aux": ([_V, _D], _i),
    "clawgpu_bc2": ([_V, _I], _i),
    "clawgpu_step2": ([_V, _d, _d, _d, _I, _I, _i, _D], _i),
    # device pointers (qold, qnew, cfl) and the stream as plain addresses
    "clawgpu_step2_noalloc": ([_i, _D, _i, _i, _i, _i, _i, C.c_long, C.c_long, _V, _V, _d, _d, _d,
                               _I, _I, _i, _V, _V], _i),
    "clawgpu_dimsp2": ([_V, _d, _d, _d, _I, _I, _D, _i, _D], _i),
    "clawgpu_claw2": ([_V, _d, _d, _d, _D, _D, _D, _I, _I, _I, _I, _i, _I], _i),
    "clawgpu_claw2_hooks": ([_V, _d, _d, _d, _d, _d, _D, _D, _D, _I, _I, _I, _I, _i, _V, _I], _i),
    "clawgpu_claw2_dev_hooks": ([_V, _d, _d, _d, _d, _d, _D, _D, _D, _I, _I, _I, _I, _i, _V, _I],
                                _i),
    "clawgpu_dev_hooks_load": ([C.POINTER(_V), C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                _i], _i),
    "clawgpu_dev_hooks_destroy": ([_V], _i),
    "clawgpu_solver2_load": ([C.POINTER(_V), C.c_char_p, C.c_char_p], _i),
    "clawgpu_solver2_destroy": ([_V], _i),
    "clawgpu_solver2_shape": ([_V, _I, _I, _I], _i),
    "clawgpu_create_user": ([C.POINTER(_V), _V, _i, _i, _i, _D, _i], _i),
    "clawgpu_dev_hooks_user": ([_V, _D, _D, _i], _i),
    "clawgpu_set_segment_rows": ([_V, _i], _i),
    "clawgpu_get_segment_rows": ([_V, _I], _i),
    "clawgpu_set_arith": ([_V, _i], _i),
    "clawgpu_solver_probe": ([_V, _i, _i, _D, _D, _D, _D, _D, _D, _D, _D, _D], _i),
    "clawgpu_learned_set_params": ([_V, C.POINTER(C.c_float), _i], _i),
    "clawgpu_learned_features": ([_V, _d, _d, _d], _i),
    "clawgpu_learned_phi": ([_V, _i], _i),
    "clawgpu_learned_table_size": ([_V, C.POINTER(C.c_long)], _i),
    "clawgpu_learned_read": ([_V, _i, _V, C.c_long], _i),
    "clawgpu_learned_write": ([_V, _i, _V, C.c_long], _i),
    "clawgpu_learned_set_target": ([_V, _D], _i),
    "clawgpu_learned_grad": ([_V, _d, _d, _d, _I, _I, _i, C.POINTER(C.c_float), _D], _i),
    "clawgpu_learned_time": ([_V, _d, _d, _d, _i, C.POINTER(C.c_float), C.POINTER(C.c_float)], _i),
    "clawgpu_time_steps": ([_V, _d, _d, _d, _I, _I, _I, _i, _i, _F, _F, _D], _i),
    "clawgpu_copy_calibration": ([C.c_long, _i, _F], _i),
    "clawgpu_copy_bandwidth": ([C.c_long, _i, _i, _F], _i),
    "clawgpu_fp64_fma_peak": ([_i, _F, _D], _i),
    "clawgpu_batch_create": ([C.POINTER(_V), _i, _i, _I, _I, _I, _D, _I, _D, _D], _i),
    "clawgpu_batch_destroy": ([_V], _i),
    "clawgpu_batch_set_q": ([_V, _i, _D], _i),
    "clawgpu_batch_get_q": ([_V, _i, _D], _i),
    "clawgpu_batch_set_arith": ([_V, _i], _i),
    "clawgpu_batch_start": ([_V, _d, _D, _D, _D, _i, _I, _I, _i], _i),
    "clawgpu_batch_advance": ([_V, _i, _i, _I], _i),
    "clawgpu_batch_result": ([_V, _i, _D, _D, _D, _I, _I], _i),
    "clawgpu_batch_set_schedule": ([_V, _I], _i),
    "clawgpu_batch_get_schedule": ([_V, _I], _i),
    "clawgpu_batch_tune": ([_V, _i, _I, _I, _F, _i], _i),
    "clawgpu_batch_time": ([_V, _i, _F, _F], _i),
    "clawgpu1_create": ([C.POINTER(_V), _i, _i, _i, _i, _i, _i, _D, _i], _i),
    "clawgpu1_destroy": ([_V], _i),
    "clawgpu1_set_q": ([_V, _D], _i),
    "clawgpu1_get_q": ([_V, _D], _i),
    "clawgpu1_set_aux": ([_V, _D], _i),
    "clawgpu1_bc1": ([_V, _I], _i),
    "clawgpu1_step1": ([_V, _d, _d, _I, _I, _i, _D], _i),
    "clawgpu1_claw1": ([_V, _d, _d, _D, _D, _D, _I, _I, _I, _I, _i, _I], _i),
    "clawgpu1_claw1_hooks": ([_V, _d, _d, _d, _D, _D, _D, _I, _I, _I, _I, _i, _V, _I], _i),
    "clawgpu3_create": ([C.POINTER(_V), _i, _i, _i, _i, _i, _i, _i, _i], _i),
    "clawgpu3_destroy": ([_V], _i),
    "clawgpu3_set_q": ([_V, _D], _i),
    "clawgpu3_get_q": ([_V, _D], _i),
    "clawgpu3_set_aux": ([_V, _D], _i),
    "clawgpu3_bc3": ([_V, _D, _d, _I], _i),
    "clawgpu3_step3": ([_V, _d, _d, _d, _d, _I, _I, _i, _D], _i),
    "clawgpu3_dimsp3": ([_V, _d, _d, _d, _d, _I, _I, _D, _i, _D], _i),
    "clawgpu3_claw3": ([_V, _D, _d, _D, _D, _D, _I, _I, _I, _I, _i, _I], _i),
    "clawgpu3_claw3_hooks": ([_V, _D, _d, _D, _D, _D, _I, _I, _I, _I, _i, _V, _I], _i),
    "clawgpu3_claw3_dev_hooks": ([_V, _D, _d, _D, _D, _D, _I, _I, _I, _I, _i, _V, _I], _i),
    "clawgpu_solver3_load": ([C.POINTER(_V), C.c_char_p, C.c_char_p], _i),
    "clawgpu_solver3_destroy": ([_V], _i),
    "clawgpu_solver3_shape": ([_V, _I, _I, _I, _I], _i),
    "clawgpu3_create_user": ([C.POINTER(_V), _V, _i, _i, _i, _i, _D, _i], _i),
    "clawgpu3_set_segment": ([_V, _i], _i),
    "clawgpu3_set_tile": ([_V, _i, _i], _i),
    "clawgpu3_set_arith": ([_V, _i], _i),
    "clawgpu3_get_tile": ([_V, _I, _I], _i),
    "clawgpu3_set_aux_table": ([_V, _i, _I], _i),
    "clawgpu3_time_steps": ([_V, _D, _d, _I, _I, _I, _i, _i, _F, _F, _D], _i),
}

_lib = None


class ClawGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"clawgpu error {code}: {msg}")
        self.code = code


def declared_symbols():
    """Names of the functions declared in include/clawgpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(clawgpu[13]?_\w+)\s*\(", text, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m classic_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().clawgpu_last_error()
        raise ClawGpuError(rc, msg.decode() if msg else "")
    return rc
