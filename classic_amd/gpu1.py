"""Python front end of the 1D device path: a context per grid.

``Grid1D`` mirrors the arguments of the reference's claw1/step1/bc1 (method(1:7),
mthlim, mthbc(1:2)).  Host arrays use the reference layout q(meqn,
1-mbc:mx+mbc): C-contiguous numpy arrays of shape (mx+2*mbc, meqn).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SOLVER_ACOUSTICS, SOLVER_ADVECTION, SOLVER_VC_ACOUSTICS, check
from .gpu import _dp, _guarded, _ip, _ivec

SHAPE1 = {SOLVER_ACOUSTICS: (2, 2), SOLVER_VC_ACOUSTICS: (2, 2), SOLVER_ADVECTION: (1, 1)}

_DBL = C.c_double
_PD = C.POINTER(_DBL)
_BC1_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, _DBL, _DBL, _PD, C.c_int, _PD,
                      _DBL, _DBL, C.POINTER(C.c_int))
_B4STEP1_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, _PD, _DBL, _DBL, _DBL, _DBL,
                          C.c_int, _PD)
_SRC1_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, _DBL, _DBL, _PD, C.c_int, _PD,
                       _DBL, _DBL)


class _Hooks1(C.Structure):
    _fields_ = [("user", C.c_void_p), ("bc1", _BC1_FN), ("b4step1", _B4STEP1_FN),
                ("src1", _SRC1_FN)]


def _view1(ptr, n, m):
    if not ptr or m == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n, m))


def make_hooks1(bc1=None, b4step1=None, src1=None):
    """C hook table for clawgpu1_claw1_hooks from Python callables, called
    with numpy views (shape (mx+2*mbc, m)) of the host copies of q and aux
    to modify in place: bc1(q, aux, t, dt, mthbc), b4step1(q, aux, t, dt),
    src1(q, aux, t, dt).  Returns (table, keep, errors) as make_hooks."""
    errors, keep, h = [], [], _Hooks1()
    if bc1 is not None:
        def _bc(user, meqn, mbc, mx, xl, dx, q, maux, aux, t, dt, mthbc):
            n = mx + 2 * mbc
            bc1(_view1(q, n, meqn), _view1(aux, n, maux), t, dt, (mthbc[0], mthbc[1]))
        keep.append(_BC1_FN(_guarded(_bc, errors)))
        h.bc1 = keep[-1]
    if b4step1 is not None:
        def _b4(user, mbc, mx, meqn, q, xl, dx, t, dt, maux, aux):
            n = mx + 2 * mbc
            b4step1(_view1(q, n, meqn), _view1(aux, n, maux), t, dt)
        keep.append(_B4STEP1_FN(_guarded(_b4, errors)))
        h.b4step1 = keep[-1]
    if src1 is not None:
        def _src(user, meqn, mbc, mx, xl, dx, q, maux, aux, t, dt):
            n = mx + 2 * mbc
            src1(_view1(q, n, meqn), _view1(aux, n, maux), t, dt)
        keep.append(_SRC1_FN(_guarded(_src, errors)))
        h.src1 = keep[-1]
    return h, keep, errors


class Grid1D:
    def __init__(self, solver_id, mx, params=(), maux=0, mbc=2):
        self.lib = _lib.load()
        self.meqn, self.mwaves = SHAPE1[solver_id]
        self.mx, self.mbc, self.maux, self.solver_id = mx, mbc, maux, solver_id
        p = np.ascontiguousarray(np.asarray(params if len(params) else [0.0], dtype=np.float64))
        h = C.c_void_p()
        check(self.lib.clawgpu1_create(C.byref(h), self.meqn, self.mwaves, maux, mbc, mx,
                                       solver_id, _dp(p), len(params)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu1_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != (self.mx + 2 * self.mbc, self.meqn):
            raise ValueError(f"q has shape {q.shape}")
        check(self.lib.clawgpu1_set_q(self.h, _dp(q)))

    def get_q(self):
        q = np.empty((self.mx + 2 * self.mbc, self.meqn))
        check(self.lib.clawgpu1_get_q(self.h, _dp(q)))
        return q

    def set_aux(self, aux):
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        if aux.shape != (self.mx + 2 * self.mbc, self.maux):
            raise ValueError(f"aux has shape {aux.shape}")
        check(self.lib.clawgpu1_set_aux(self.h, _dp(aux)))

    def bc1(self, mthbc):
        check(self.lib.clawgpu1_bc1(self.h, _ip(_ivec(mthbc, 2))))

    def step1(self, dx, dt, method=(1, 2, 0, 0, 0, 0, 0), mthlim=(4, 4), use_fwave=False):
        cfl = C.c_double(0.0)
        check(self.lib.clawgpu1_step1(self.h, dx, dt, _ip(_ivec(method, 7)),
                                      _ip(_ivec(mthlim, self.mwaves)), int(use_fwave),
                                      C.byref(cfl)))
        return cfl.value

    def claw1(self, dx, tstart, tend, dtv, cflv, nmax, method=(1, 2, 0, 0, 0, 0, 0),
              mthlim=(4, 4), mthbc=(1, 1), use_fwave=False, bc1=None, b4step1=None, src1=None,
              xlower=0.0):
        """claw1 time loop on the device q; returns a dict of the outputs.
        bc1 / b4step1 / src1: optional user routines (see make_hooks1)."""
        te = C.c_double(tend)
        dtv_a = np.zeros(5)
        dtv_a[: len(dtv)] = dtv
        cflv_a = np.zeros(4)
        cflv_a[: len(cflv)] = cflv
        nv = np.array([nmax, 0], dtype=np.int32)
        info = C.c_int(0)
        args = (_dp(dtv_a), _dp(cflv_a), _ip(nv), _ip(_ivec(method, 7)),
                _ip(_ivec(mthlim, self.mwaves)), _ip(_ivec(mthbc, 2)), int(use_fwave))
        if bc1 is None and b4step1 is None and src1 is None:
            check(self.lib.clawgpu1_claw1(self.h, dx, tstart, C.byref(te), *args, C.byref(info)))
        else:
            hooks, keep, errors = make_hooks1(bc1, b4step1, src1)
            rc = self.lib.clawgpu1_claw1_hooks(self.h, dx, float(xlower), tstart, C.byref(te), *args,
                                               C.byref(hooks), C.byref(info))
            del keep
            if errors:
                raise errors[0]
            check(rc)
        return dict(info=info.value, tend=te.value, dtv=dtv_a, cflv=cflv_a, nv=nv)

    def frames(self, dx, t0, tfinal, nout, dt_initial, cflv=(1.0, 0.9), dt_max=1e99,
               nmax=50000, **kw):
        out = []
        dtv = [dt_initial, dt_max]
        tend = t0
        dtout = (tfinal - t0) / float(nout)
        for _ in range(nout):
            tstart = tend
            r = self.claw1(dx, tstart, tstart + dtout, dtv, cflv, nmax, **kw)
            if r["info"] != 0:
                raise RuntimeError(f"claw1 info={r['info']}")
            dtv = [r["dtv"][4], dt_max]
            tend = r["tend"]
            out.append((tend, self.get_q()))
        return out
