"""Python front end of the 3D device path: a context per grid.

``Grid3D`` mirrors the arguments of the reference's claw3/step3/dimsp3/bc3
(sizes, method(1:7), mthlim, mthbc(1:6)) and keeps q and aux resident on the
GPU.  Host arrays use the reference layout q(meqn, 1-mbc:mx+mbc,
1-mbc:my+mbc, 1-mbc:mz+mbc), i.e. C-contiguous numpy arrays of shape
(mz+2*mbc, my+2*mbc, mx+2*mbc, meqn); aux likewise with maux last.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SOLVER_VC_ACOUSTICS, check
from .gpu import _dp, _guarded, _ip, _ivec

_DBL = C.c_double
_PD = C.POINTER(_DBL)
_I = C.c_int
_BC3_FN = C.CFUNCTYPE(None, C.c_void_p, _I, _I, _I, _I, _I, _DBL, _DBL, _DBL, _DBL, _DBL, _DBL, _PD,
                      _I, _PD, _DBL, _DBL, C.POINTER(_I))
_B4STEP3_FN = C.CFUNCTYPE(None, C.c_void_p, _I, _I, _I, _I, _I, _PD, _DBL, _DBL, _DBL, _DBL, _DBL,
                          _DBL, _DBL, _DBL, _I, _PD)
_SRC3_FN = C.CFUNCTYPE(None, C.c_void_p, _I, _I, _I, _I, _I, _DBL, _DBL, _DBL, _DBL, _DBL, _DBL,
                       _PD, _I, _PD, _DBL, _DBL)


class _Hooks3(C.Structure):
    _fields_ = [("user", C.c_void_p), ("bc3", _BC3_FN), ("b4step3", _B4STEP3_FN),
                ("src3", _SRC3_FN)]


def _view3(ptr, mbc, mx, my, mz, m):
    if not ptr or m == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(mz + 2 * mbc, my + 2 * mbc, mx + 2 * mbc, m))


def make_hooks3(bc3=None, b4step3=None, src3=None):
    """C hook table for clawgpu3_claw3_hooks from Python callables, called
    with numpy views (shape (mz+2*mbc, my+2*mbc, mx+2*mbc, m)) of the host
    copies of q and aux to modify in place: bc3(q, aux, t, dt, mthbc),
    b4step3(q, aux, t, dt), src3(q, aux, t, dt).  Returns (table, keep,
    errors) as gpu.make_hooks."""
    errors, keep, h = [], [], _Hooks3()
    if bc3 is not None:
        def _bc(user, meqn, mbc, mx, my, mz, xl, yl, zl, dx, dy, dz, q, maux, aux, t, dt, mthbc):
            bc3(_view3(q, mbc, mx, my, mz, meqn), _view3(aux, mbc, mx, my, mz, maux), t, dt,
                tuple(mthbc[k] for k in range(6)))
        keep.append(_BC3_FN(_guarded(_bc, errors)))
        h.bc3 = keep[-1]
    if b4step3 is not None:
        def _b4(user, mbc, mx, my, mz, meqn, q, xl, yl, zl, dx, dy, dz, t, dt, maux, aux):
            b4step3(_view3(q, mbc, mx, my, mz, meqn), _view3(aux, mbc, mx, my, mz, maux), t, dt)
        keep.append(_B4STEP3_FN(_guarded(_b4, errors)))
        h.b4step3 = keep[-1]
    if src3 is not None:
        def _src(user, meqn, mbc, mx, my, mz, xl, yl, zl, dx, dy, dz, q, maux, aux, t, dt):
            src3(_view3(q, mbc, mx, my, mz, meqn), _view3(aux, mbc, mx, my, mz, maux), t, dt)
        keep.append(_SRC3_FN(_guarded(_src, errors)))
        h.src3 = keep[-1]
    return h, keep, errors


class Claw3Result:
    def __init__(self, info, tend, dtv, cflv, nv):
        self.info, self.tend, self.dtv, self.cflv, self.nv = info, tend, dtv, cflv, nv


class UserSolver3:
    """A 3D user Riemann solver: solver `name` of the gfx950 code object at
    path, built with CLAWGPU_EXPORT_SOLVER3 (include/clawgpu_solver.h).  Pass
    it to Grid3D as solver; keep it open while its grids exist."""

    def __init__(self, path, name):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.clawgpu_solver3_load(C.byref(h), str(path).encode(), name.encode()))
        self.h = h
        m, w, a, r = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(self.lib.clawgpu_solver3_shape(h, C.byref(m), C.byref(w), C.byref(a), C.byref(r)))
        self.meqn, self.mwaves, self.maux, self.arith = m.value, w.value, a.value, r.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu_solver3_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Grid3D:
    """Device-resident 3D grid: variable-coefficient acoustics (q = (p, u, v,
    w)), or a UserSolver3 (solver=..., params reach it as SolverParams.user;
    the split step only)."""

    def __init__(self, mx, my, mz, maux=2, mbc=2, lower=(-1.0, -1.0, -1.0),
                 upper=(1.0, 1.0, 1.0), solver=None, params=()):
        self.lib = _lib.load()
        self.mx, self.my, self.mz, self.mbc, self.maux = mx, my, mz, mbc, maux
        d = [(upper[a] - lower[a]) / n for a, n in enumerate((mx, my, mz))]
        self.geom = np.array([lower[0], lower[1], lower[2], d[0], d[1], d[2]], dtype=np.float64)
        h = C.c_void_p()
        self.solver = solver
        if solver is not None:
            self.meqn, self.mwaves = solver.meqn, solver.mwaves
            p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
            check(self.lib.clawgpu3_create_user(C.byref(h), solver.h, maux, mx, my, mz, _dp(p),
                                                p.size))
        else:
            self.meqn, self.mwaves = 4, 2
            check(self.lib.clawgpu3_create(C.byref(h), 4, 2, maux, mbc, mx, my, mz,
                                           SOLVER_VC_ACOUSTICS))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu3_destroy(self.h)
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

    @property
    def dx(self):
        return tuple(self.geom[3:])

    def shape(self, ncomp=None):
        return (self.mz + 2 * self.mbc, self.my + 2 * self.mbc, self.mx + 2 * self.mbc,
                self.meqn if ncomp is None else ncomp)

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.shape():
            raise ValueError(f"q has shape {q.shape}, expected {self.shape()}")
        check(self.lib.clawgpu3_set_q(self.h, _dp(q)))

    def get_q(self):
        q = np.empty(self.shape())
        check(self.lib.clawgpu3_get_q(self.h, _dp(q)))
        return q

    def set_aux(self, aux):
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        if aux.shape != self.shape(self.maux):
            raise ValueError(f"aux has shape {aux.shape}, expected {self.shape(self.maux)}")
        check(self.lib.clawgpu3_set_aux(self.h, _dp(aux)))

    def bc3(self, mthbc, t=0.0):
        check(self.lib.clawgpu3_bc3(self.h, _dp(self.geom), t, _ip(_ivec(mthbc, 6))))

    def step3(self, dt, method=(1, 2, 22, 0, 0, 0, 2), mthlim=(3, 3), use_fwave=False):
        cfl = C.c_double(0.0)
        dx, dy, dz = self.dx
        check(self.lib.clawgpu3_step3(self.h, dx, dy, dz, dt, _ip(_ivec(method, 7)),
                                      _ip(_ivec(mthlim, self.mwaves)), int(use_fwave), C.byref(cfl)))
        return cfl.value

    def dimsp3(self, dt, method=(1, 2, -1, 0, 0, 0, 2), mthlim=(3, 3), cflv=(1.0, 0.9),
               use_fwave=False):
        cfl = C.c_double(0.0)
        cv = np.zeros(4)
        cv[: len(cflv)] = cflv
        dx, dy, dz = self.dx
        check(self.lib.clawgpu3_dimsp3(self.h, dx, dy, dz, dt, _ip(_ivec(method, 7)),
                                       _ip(_ivec(mthlim, self.mwaves)), _dp(cv), int(use_fwave),
                                       C.byref(cfl)))
        return cfl.value

    def claw3(self, tstart, tend, dtv, cflv, nmax, method=(1, 2, 22, 0, 0, 0, 2), mthlim=(3, 3),
              mthbc=(1, 1, 1, 1, 1, 1), use_fwave=False, bc3=None, b4step3=None, src3=None,
              dev_hooks=None):
        """The claw3 time loop on the device; returns Claw3Result.
        bc3 / b4step3 / src3: optional user routines (see make_hooks3);
        dev_hooks: a gpu.DeviceHooks of 3D device routines instead."""
        te = C.c_double(tend)
        dtv_a = np.zeros(5)
        dtv_a[: len(dtv)] = dtv
        cflv_a = np.zeros(4)
        cflv_a[: len(cflv)] = cflv
        nv = np.array([nmax, 0], dtype=np.int32)
        info = C.c_int(0)
        args = (_dp(self.geom), tstart, C.byref(te), _dp(dtv_a), _dp(cflv_a), _ip(nv),
                _ip(_ivec(method, 7)), _ip(_ivec(mthlim, self.mwaves)), _ip(_ivec(mthbc, 6)),
                int(use_fwave))
        if dev_hooks is not None:
            check(self.lib.clawgpu3_claw3_dev_hooks(self.h, *args, dev_hooks.h, C.byref(info)))
        elif bc3 is None and b4step3 is None and src3 is None:
            check(self.lib.clawgpu3_claw3(self.h, *args, C.byref(info)))
        else:
            hooks, keep, errors = make_hooks3(bc3, b4step3, src3)
            rc = self.lib.clawgpu3_claw3_hooks(self.h, *args, C.byref(hooks), C.byref(info))
            del keep
            if errors:
                raise errors[0]
            check(rc)
        return Claw3Result(info.value, te.value, dtv_a, cflv_a, nv)

    def frames(self, t0, tfinal, nout, dt_initial, cflv=(1.0, 0.9), dt_max=1e9, nmax=50000,
               **kw):
        """claw3ez.f's output_style 1 frame loop; returns [(t, q)]."""
        out = []
        dtv = [dt_initial, dt_max]
        tend = t0
        dtout = (tfinal - t0) / float(nout)
        for _ in range(nout):
            tstart = tend
            r = self.claw3(tstart, tstart + dtout, dtv, cflv, nmax, **kw)
            if r.info != 0:
                raise RuntimeError(f"claw3 info={r.info}")
            dtv = [r.dtv[4], dt_max]
            tend = r.tend
            out.append((tend, self.get_q()))
        return out

    def set_arith(self, mode):
        """Unsplit step arithmetic: 1 fast (default), 0 exact (the reference's
        rounding and order, bit for bit; steps with a capacity function
        always run exact)."""
        check(self.lib.clawgpu3_set_arith(self.h, int(mode)))

    def set_segment(self, cells):
        check(self.lib.clawgpu3_set_segment(self.h, cells))

    def set_aux_table(self, enable=True):
        """Read the material table (when set_aux found at most 64 distinct
        (Z, c) pairs) or aux itself; returns the number of materials found."""
        n = C.c_int(0)
        check(self.lib.clawgpu3_set_aux_table(self.h, int(bool(enable)), C.byref(n)))
        return n.value

    def set_tile(self, tile, xcd_order=2):
        """Tile shape of the unsplit kernel (0..4: tile kernels of 4x4, 4x2,
        2x2, 8x2, 2x4 y-z cells; 5, 6: z-marching kernel with 10 or 6 rows
        per block, which needs the material table and otherwise runs shape
        4) and tile order: order bit 0 gives each XCD a contiguous range of
        tiles, bit 1 numbers tiles with z fastest (tile kernels only)."""
        check(self.lib.clawgpu3_set_tile(self.h, tile, int(xcd_order)))

    def get_tile(self):
        """(tile shape, tile order) of the unsplit kernel."""
        t, o = C.c_int(0), C.c_int(0)
        check(self.lib.clawgpu3_get_tile(self.h, C.byref(t), C.byref(o)))
        return t.value, o.value

    def time_steps(self, dt, nsteps, method=(1, 2, 22, 0, 0, 0, 2), mthlim=(3, 3),
                   mthbc=(3, 3, 3, 3, 3, 3), use_fwave=False):
        tot = C.c_float(0.0)
        ker = C.c_float(0.0)
        cfl = C.c_double(0.0)
        check(self.lib.clawgpu3_time_steps(self.h, _dp(self.geom), dt, _ip(_ivec(method, 7)),
                                           _ip(_ivec(mthlim, self.mwaves)), _ip(_ivec(mthbc, 6)),
                                           int(use_fwave), nsteps, C.byref(tot), C.byref(ker),
                                           C.byref(cfl)))
        return tot.value, ker.value, cfl.value
