"""Python front end of the device step: a context per grid.

``Grid2D`` mirrors the arguments of the reference's claw2/step2 (sizes,
method(1:7), mthlim, mthbc) and keeps q resident on the GPU.  Host arrays use
the reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc), i.e. C-contiguous
numpy arrays of shape (my+2*mbc, mx+2*mbc, meqn).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SOLVER_ACOUSTICS, SOLVER_EULER, SOLVER_VC_ADVECTION, check

SOLVER_SHAPE = {SOLVER_ACOUSTICS: (3, 2), SOLVER_EULER: (4, 3), SOLVER_VC_ADVECTION: (1, 1)}


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ivec(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} integers, got {a.size}")
    return a


def solver_params(solver_id, rho=1.0, bulk=4.0, gamma=1.4, efix=True):
    if solver_id == SOLVER_ACOUSTICS:
        return [float(rho), float(bulk)]
    if solver_id == SOLVER_EULER:
        return [float(gamma), 1.0 if efix else 0.0]
    raise ValueError("unknown solver")


_DBL = C.c_double
_BC2_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _DBL, _DBL, _DBL, _DBL,
                      C.POINTER(_DBL), C.c_int, C.POINTER(_DBL), _DBL, _DBL, C.POINTER(C.c_int))
_B4STEP2_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_DBL),
                          _DBL, _DBL, _DBL, _DBL, _DBL, _DBL, C.c_int, C.POINTER(_DBL))
_SRC2_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _DBL, _DBL, _DBL, _DBL,
                       C.POINTER(_DBL), C.c_int, C.POINTER(_DBL), _DBL, _DBL)


class _Hooks(C.Structure):
    _fields_ = [("user", C.c_void_p), ("bc2", _BC2_FN), ("b4step2", _B4STEP2_FN),
                ("src2", _SRC2_FN)]


def _view(ptr, my, mx, m):
    if not ptr or m == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(my, mx, m))


def _guarded(fn, errors):
    """fn wrapped for use as a C callback: an exception cannot propagate
    through C, so it is appended to errors and re-raised by the caller."""
    def run(*a):
        try:
            fn(*a)
        except BaseException as e:
            errors.append(e)
    return run


def make_hooks(bc2=None, b4step2=None, src2=None):
    """C hook table for clawgpu_claw2_hooks from Python callables, each
    called with numpy views of the host copies of q and aux (reference
    layout, ghost cells included; aux None when maux = 0) to modify in place:
    bc2(q, aux, t, dt, mthbc), b4step2(q, aux, t, dt), src2(q, aux, t, dt).
    Returns (table, keep, errors): keep holds the ctypes callbacks alive,
    errors collects exceptions raised inside the hooks."""
    errors = []
    h = _Hooks()
    keep = []

    def guard(fn):
        return _guarded(fn, errors)

    if bc2 is not None:
        def _bc(user, meqn, mbc, mx, my, xl, yl, dx, dy, q, maux, aux, t, dt, mthbc):
            n = (my + 2 * mbc, mx + 2 * mbc)
            bc2(_view(q, *n, meqn), _view(aux, *n, maux), t, dt, tuple(mthbc[k] for k in range(4)))
        f = _BC2_FN(guard(_bc))
        keep.append(f)
        h.bc2 = f
    if b4step2 is not None:
        def _b4(user, mbc, mx, my, meqn, q, xl, yl, dx, dy, t, dt, maux, aux):
            n = (my + 2 * mbc, mx + 2 * mbc)
            b4step2(_view(q, *n, meqn), _view(aux, *n, maux), t, dt)
        f = _B4STEP2_FN(guard(_b4))
        keep.append(f)
        h.b4step2 = f
    if src2 is not None:
        def _src(user, meqn, mbc, mx, my, xl, yl, dx, dy, q, maux, aux, t, dt):
            n = (my + 2 * mbc, mx + 2 * mbc)
            src2(_view(q, *n, meqn), _view(aux, *n, maux), t, dt)
        f = _SRC2_FN(guard(_src))
        keep.append(f)
        h.src2 = f
    return h, keep, errors


class DeviceHooks:
    """Device-side user routines: the kernels named bc2 / b4step2 / src2 in
    the gfx950 code object at path (see include/clawgpu.h,
    clawgpu_dev_grid), with nuser doubles of device memory passed to them."""

    def __init__(self, path, bc2=None, b4step2=None, src2=None, nuser=0):
        self.lib = _lib.load()
        self.nuser = nuser
        h = C.c_void_p()
        enc = (lambda x: None if x is None else x.encode())
        check(self.lib.clawgpu_dev_hooks_load(C.byref(h), str(path).encode(), enc(bc2),
                                              enc(b4step2), enc(src2), int(nuser)))
        self.h = h

    def write_user(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(self.lib.clawgpu_dev_hooks_user(self.h, _dp(v), None, v.size))

    def read_user(self, n=None):
        out = np.empty(self.nuser if n is None else n)
        check(self.lib.clawgpu_dev_hooks_user(self.h, None, _dp(out), out.size))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu_dev_hooks_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class UserSolver:
    """A user Riemann solver: solver `name` of the gfx950 code object at path,
    built from include/clawgpu_solver.h.  Pass it to Grid2D in place of a
    solver id; keep it open while its grids exist."""

    def __init__(self, path, name):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.clawgpu_solver2_load(C.byref(h), str(path).encode(), name.encode()))
        self.h = h
        m, w, a = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.clawgpu_solver2_shape(h, C.byref(m), C.byref(w), C.byref(a)))
        self.meqn, self.mwaves, self.arith = m.value, w.value, a.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu_solver2_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Grid2D:
    """Device-resident single grid (one clawgpu context).  solver_id is a
    built-in solver id or a UserSolver (whose params reach it as
    SolverParams.user)."""

    def __init__(self, solver_id, mx, my, params, mbc=2, maux=0, arith=0):
        self.lib = _lib.load()
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        h = C.c_void_p()
        if isinstance(solver_id, UserSolver):
            meqn, mwaves = solver_id.meqn, solver_id.mwaves
            check(self.lib.clawgpu_create_user(C.byref(h), solver_id.h, mbc, mx, my, _dp(p), p.size))
            self.solver = solver_id
            solver_id = _lib.SOLVER_USER
        else:
            meqn, mwaves = SOLVER_SHAPE[solver_id]
            check(self.lib.clawgpu_create(C.byref(h), meqn, mwaves, maux, mbc, mx, my, solver_id,
                                          _dp(p), p.size))
        self.solver_id, self.mx, self.my, self.mbc = solver_id, mx, my, mbc
        self.meqn, self.mwaves, self.maux = meqn, mwaves, maux
        self.h = h
        if arith:
            self.set_arith(arith)

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.my + 2 * self.mbc, self.mx + 2 * self.mbc, self.meqn)

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.shape:
            raise ValueError(f"q has shape {q.shape}, expected {self.shape}")
        check(self.lib.clawgpu_set_q(self.h, _dp(q)))

    def set_aux(self, aux):
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        want = self.shape[:2] + (self.maux,)
        if aux.shape != want:
            raise ValueError(f"aux has shape {aux.shape}, expected {want}")
        check(self.lib.clawgpu_set_aux(self.h, _dp(aux)))

    def get_q(self):
        q = np.empty(self.shape)
        check(self.lib.clawgpu_get_q(self.h, _dp(q)))
        return q

    def bc2(self, mthbc):
        check(self.lib.clawgpu_bc2(self.h, _ip(_ivec(mthbc, 4))))

    def step2(self, dx, dy, dt, method, mthlim, use_fwave=False):
        """One unsplit step on the device q (ghosts already set); returns cfl."""
        cfl = C.c_double(0.0)
        check(self.lib.clawgpu_step2(self.h, dx, dy, dt, _ip(_ivec(method, 7)),
                                     _ip(_ivec(mthlim, self.mwaves)), int(bool(use_fwave)),
                                     C.byref(cfl)))
        return cfl.value

    def dimsp2(self, dx, dy, dt, method, mthlim, cflv=(1.0, 0.9), use_fwave=False):
        """One dimensionally split step (method(3) = -1 or -2); returns cfl."""
        cfl = C.c_double(0.0)
        cv = np.ascontiguousarray(np.asarray(cflv, dtype=np.float64))
        check(self.lib.clawgpu_dimsp2(self.h, dx, dy, dt, _ip(_ivec(method, 7)),
                                      _ip(_ivec(mthlim, self.mwaves)), _dp(cv),
                                      int(bool(use_fwave)), C.byref(cfl)))
        return cfl.value

    def claw2(self, dx, dy, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc,
              use_fwave=False, bc2=None, b4step2=None, src2=None, xlower=0.0, ylower=0.0,
              dev_hooks=None):
        """claw2 time loop on the device q; returns dict of the outputs.
        bc2 / b4step2 / src2: optional user routines (see make_hooks);
        dev_hooks: a DeviceHooks of device routines instead."""
        if dev_hooks is not None:
            return self._claw2_hooks(dx, dy, tstart, tend, dtv, cflv, nmax, method, mthlim,
                                     mthbc, use_fwave, None, None, None, xlower, ylower,
                                     dev_hooks)
        if bc2 is not None or b4step2 is not None or src2 is not None:
            return self._claw2_hooks(dx, dy, tstart, tend, dtv, cflv, nmax, method, mthlim,
                                     mthbc, use_fwave, bc2, b4step2, src2, xlower, ylower)
        te = C.c_double(tend)
        dtv_a = np.zeros(5)
        dtv_a[: len(dtv)] = dtv
        cflv_a = np.zeros(4)
        cflv_a[: len(cflv)] = cflv
        nv = np.array([nmax, 0], dtype=np.int32)
        info = C.c_int(0)
        m = _ivec(method, 7).copy()
        m[6] = self.maux
        rc = self.lib.clawgpu_claw2(self.h, dx, dy, tstart, C.byref(te), _dp(dtv_a), _dp(cflv_a),
                                    _ip(nv), _ip(m), _ip(_ivec(mthlim, self.mwaves)),
                                    _ip(_ivec(mthbc, 4)), int(bool(use_fwave)), C.byref(info))
        check(rc)
        return dict(info=info.value, tend=te.value, dtv=dtv_a, cflv=cflv_a, nv=nv)

    def _claw2_hooks(self, dx, dy, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc,
                     use_fwave, bc2, b4step2, src2, xlower, ylower, dev_hooks=None):
        te = C.c_double(tend)
        dtv_a = np.zeros(5)
        dtv_a[: len(dtv)] = dtv
        cflv_a = np.zeros(4)
        cflv_a[: len(cflv)] = cflv
        nv = np.array([nmax, 0], dtype=np.int32)
        info = C.c_int(0)
        m = _ivec(method, 7).copy()
        m[6] = self.maux
        if dev_hooks is not None:
            rc = self.lib.clawgpu_claw2_dev_hooks(
                self.h, dx, dy, float(xlower), float(ylower), tstart, C.byref(te), _dp(dtv_a),
                _dp(cflv_a), _ip(nv), _ip(m), _ip(_ivec(mthlim, self.mwaves)),
                _ip(_ivec(mthbc, 4)), int(bool(use_fwave)), dev_hooks.h, C.byref(info))
            check(rc)
            return dict(info=info.value, tend=te.value, dtv=dtv_a, cflv=cflv_a, nv=nv)
        hooks, keep, errors = make_hooks(bc2, b4step2, src2)
        rc = self.lib.clawgpu_claw2_hooks(self.h, dx, dy, float(xlower), float(ylower), tstart,
                                          C.byref(te), _dp(dtv_a), _dp(cflv_a), _ip(nv), _ip(m),
                                          _ip(_ivec(mthlim, self.mwaves)), _ip(_ivec(mthbc, 4)),
                                          int(bool(use_fwave)), C.byref(hooks), C.byref(info))
        del keep
        if errors:
            raise errors[0]
        check(rc)
        return dict(info=info.value, tend=te.value, dtv=dtv_a, cflv=cflv_a, nv=nv)

    def solver_probe(self, ixy, ql, qr, asdq):
        """Device rpn2/rpt2 on interfaces ql[k] | qr[k] (arrays (n, meqn));
        returns wave (n, mwaves, meqn), s (n, mwaves), amdq, apdq, bmasdq,
        bpasdq (n, meqn)."""
        ql = np.ascontiguousarray(ql, dtype=np.float64)
        qr = np.ascontiguousarray(qr, dtype=np.float64)
        asdq = np.ascontiguousarray(asdq, dtype=np.float64)
        n = ql.shape[0]
        wave = np.empty((n, self.mwaves, self.meqn))
        s = np.empty((n, self.mwaves))
        out = [np.empty((n, self.meqn)) for _ in range(4)]
        check(self.lib.clawgpu_solver_probe(self.h, ixy, n, _dp(ql), _dp(qr), _dp(asdq), _dp(wave),
                                            _dp(s), *[_dp(o) for o in out]))
        return (wave, s, *out)

    def set_segment_rows(self, rows):
        """Rows per wavefront segment of the step kernel (0: chosen per grid)."""
        check(self.lib.clawgpu_set_segment_rows(self.h, int(rows)))

    def segment_rows(self):
        """Rows per segment the last unsplit step used (0 before the first)."""
        r = C.c_int(0)
        check(self.lib.clawgpu_get_segment_rows(self.h, C.byref(r)))
        return r.value

    def set_arith(self, mode):
        """0 = exact reference arithmetic (default), 1 = fast (FMA)."""
        check(self.lib.clawgpu_set_arith(self.h, int(mode)))

    def time_steps(self, dx, dy, dt, method, mthlim, mthbc, nsteps, use_fwave=False):
        """nsteps fixed-dt steps timed with HIP events on the library's
        stream; returns (total_ms, step_kernel_ms, cfl_max)."""
        ms = C.c_float(0.0)
        kms = C.c_float(0.0)
        cfl = C.c_double(0.0)
        check(self.lib.clawgpu_time_steps(self.h, dx, dy, dt, _ip(_ivec(method, 7)),
                                          _ip(_ivec(mthlim, self.mwaves)), _ip(_ivec(mthbc, 4)),
                                          int(bool(use_fwave)), int(nsteps), C.byref(ms),
                                          C.byref(kms), C.byref(cfl)))
        return ms.value, kms.value, cfl.value


def step2(solver_id, params, qold, dx, dy, dt, method, mthlim, use_fwave=False, mbc=2,
          arith=0):
    """One step with the reference step2 contract on host arrays: returns
    (qnew, cfl).  qold must have its ghost cells set; ghost cells of qnew are
    copied from qold (the reference leaves them partially updated)."""
    qold = np.ascontiguousarray(qold, dtype=np.float64)
    my2, mx2, meqn = qold.shape
    with Grid2D(solver_id, mx2 - 2 * mbc, my2 - 2 * mbc, params, mbc=mbc, arith=arith) as g:
        g.set_q(qold)
        cfl = g.step2(dx, dy, dt, method, mthlim, use_fwave)
        qn = g.get_q()
    out = qold.copy()
    out[mbc:-mbc, mbc:-mbc] = qn[mbc:-mbc, mbc:-mbc]
    return out, cfl
