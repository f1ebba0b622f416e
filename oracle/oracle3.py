"""ctypes front end of the 3D CPU ORACLE -- test infrastructure only.

Loads ``oracle/_build/liboracle3d.so`` (the plain-C restatement in
``claw3d_oracle.c``) and, when it was built, ``oracle/_ref/libclawref3d.so``
(the reference Fortran from clawpack/classic ``src/3d`` linked with the same
variable-coefficient acoustics solvers).  Only tests and the smoke check
import this module; the product package never does.

Arrays use the reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc,
1-mbc:mz+mbc) in Fortran order, i.e. a C-contiguous numpy array of shape
(mz+2mbc, my+2mbc, mx+2mbc, meqn); aux likewise with maux last.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle3d.so")
REF_SO = os.path.join(HERE, "_ref", "libclawref3d.so")
ORACLE_OMP_SO = os.path.join(HERE, "_build", "liboracle3d_omp.so")

SOLVER_VC_ACOUSTICS = 3

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class Geom(C.Structure):
    _fields_ = [("xlower", C.c_double), ("ylower", C.c_double), ("zlower", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I)


def _arr_i(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _lib = C.CDLL(ORACLE_SO)
    return _lib


_omp = None


def lib_omp(threads: int):
    """The OpenMP build of the same restatement (-O3; the blocks of each pass
    of the unsplit step run on `threads` threads, each cell updated in the
    serial order, so results are the plain build's bit for bit)."""
    global _omp
    if _omp is None:
        if not os.path.exists(ORACLE_OMP_SO):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _omp = C.CDLL(ORACLE_OMP_SO)
    _omp.orc_set_threads(int(threads))
    return _omp


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
    return _ref


class Problem3:
    """Sizes, method flags and geometry of one 3D run (reference argument set)."""

    def __init__(self, mx, my, mz, maux=2, meqn=4, mwaves=2, mbc=2,
                 method=(1, 2, 22, 0, 0, 0, 0), mthlim=(3, 3), mthbc=(1, 1, 1, 1, 1, 1),
                 lower=(-1.0, -1.0, -1.0), upper=(1.0, 1.0, 1.0), use_fwave=False,
                 solver_id=SOLVER_VC_ACOUSTICS):
        self.mx, self.my, self.mz = mx, my, mz
        self.meqn, self.mwaves, self.maux, self.mbc = meqn, mwaves, maux, mbc
        m = list(method)
        m[6] = maux
        self.method = _arr_i(m)
        self.mthlim = _arr_i(mthlim)
        self.mthbc = _arr_i(mthbc)
        self.use_fwave = 1 if use_fwave else 0
        self.solver_id = solver_id
        d = [(upper[a] - lower[a]) / n for a, n in enumerate((mx, my, mz))]
        self.geom = Geom(lower[0], lower[1], lower[2], d[0], d[1], d[2])

    @property
    def piston(self):
        return int(self.mthbc[5] == 0)

    def shape(self, ncomp=None):
        return (self.mz + 2 * self.mbc, self.my + 2 * self.mbc, self.mx + 2 * self.mbc,
                self.meqn if ncomp is None else ncomp)


def _prep(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def step3(p: Problem3, qold, aux, dt, backend="oracle", threads=1):
    """One unsplit (method(3) >= 0) step; returns (qnew, cfl).  backend:
    "oracle", "omp" (the OpenMP build on `threads` threads) or "ref"."""
    qold = _prep(qold)
    aux = _prep(aux)
    qnew = qold.copy()
    cfl = C.c_double(0.0)
    g = p.geom
    if backend in ("oracle", "omp"):
        L = lib() if backend == "oracle" else lib_omp(threads)
        rc = L.orc3_step3(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, _dp(qold),
                              _dp(qnew), _dp(aux), C.c_double(g.dx), C.c_double(g.dy),
                              C.c_double(g.dz), C.c_double(dt), _ip(p.method), _ip(p.mthlim),
                              p.use_fwave, p.solver_id, C.byref(cfl))
    else:
        rc = ref().ref3_step3(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, _dp(qold),
                              _dp(qnew), _dp(aux), C.c_double(g.dx), C.c_double(g.dy),
                              C.c_double(g.dz), C.c_double(dt), _ip(p.method), _ip(p.mthlim),
                              p.use_fwave, C.byref(cfl))
    if rc != 0:
        raise RuntimeError(f"step3 ({backend}) failed rc={rc}")
    return qnew, cfl.value


def dimsp3(p: Problem3, qold, aux, dt, cflv=(1.0, 0.9), backend="oracle"):
    qold = _prep(qold)
    aux = _prep(aux)
    qnew = qold.copy()
    cfl = C.c_double(0.0)
    cv = np.zeros(4)
    cv[: len(cflv)] = cflv
    g = p.geom
    args = (p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, _dp(qold), _dp(qnew), _dp(aux),
            C.c_double(g.dx), C.c_double(g.dy), C.c_double(g.dz), C.c_double(dt), _ip(p.method),
            _ip(p.mthlim), _dp(cv), p.use_fwave)
    if backend == "oracle":
        rc = lib().orc3_dimsp3(*args, p.solver_id, C.byref(cfl))
    else:
        rc = ref().ref3_dimsp3(*args, C.byref(cfl))
    if rc != 0:
        raise RuntimeError(f"dimsp3 ({backend}) failed rc={rc}")
    return qnew, cfl.value


def bc3(p: Problem3, q, t=0.0, backend="oracle"):
    q = _prep(q).copy()
    if backend == "oracle":
        rc = lib().orc3_bc3(p.meqn, p.mbc, p.mx, p.my, p.mz, C.byref(p.geom), _dp(q),
                            C.c_double(t), _ip(p.mthbc))
    else:
        rc = ref().ref3_bc3(p.meqn, p.mbc, p.mx, p.my, p.mz, C.byref(p.geom), _dp(q),
                            C.c_double(t), _ip(p.mthbc), p.piston)
    if rc != 0:
        raise RuntimeError("bc3 failed")
    return q


class ClawResult:
    def __init__(self, q, info, tend, dtv, cflv, nv, dt_trace=None, cfl_trace=None):
        self.q, self.info, self.tend = q, info, tend
        self.dtv, self.cflv, self.nv = dtv, cflv, nv
        self.dt_trace, self.cfl_trace = dt_trace, cfl_trace


def claw3(p: Problem3, q, aux, tstart, tend, dtv, cflv, nmax, backend="oracle",
          trace_cap=100000):
    """The claw3 time loop (src/3d/claw3.f:596-904).  Returns ClawResult."""
    q = _prep(q).copy()
    aux = _prep(aux)
    te = C.c_double(tend)
    dtv_a = np.zeros(5)
    dtv_a[: len(dtv)] = dtv
    cflv_a = np.zeros(4)
    cflv_a[: len(cflv)] = cflv
    nv = np.array([nmax, 0], dtype=np.int32)
    if backend == "oracle":
        dtt = np.zeros(trace_cap)
        cft = np.zeros(trace_cap)
        ntr = C.c_int(0)
        info = lib().orc3_claw3(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, _dp(q),
                                _dp(aux), C.byref(p.geom), C.c_double(tstart), C.byref(te),
                                _dp(dtv_a), _dp(cflv_a), _ip(nv), _ip(p.method), _ip(p.mthlim),
                                _ip(p.mthbc), p.use_fwave, p.solver_id, _dp(dtt), _dp(cft),
                                trace_cap, C.byref(ntr))
        return ClawResult(q, info, te.value, dtv_a, cflv_a, nv, dtt[: ntr.value],
                          cft[: ntr.value])
    info = ref().ref3_claw3(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, _dp(q), _dp(aux),
                            C.byref(p.geom), C.c_double(tstart), C.byref(te), _dp(dtv_a),
                            _dp(cflv_a), _ip(nv), _ip(p.method), _ip(p.mthlim), _ip(p.mthbc),
                            p.use_fwave, p.piston)
    return ClawResult(q, info, te.value, dtv_a, cflv_a, nv)


def frames(p: Problem3, q, aux, t0, tfinal, nout, dt_initial, cflv=(1.0, 0.9), dt_max=1e9,
           nmax=50000, backend="oracle"):
    """claw3ez.f's output_style 1 frame loop: nout calls of claw3, each
    starting from the previous call's last dt.  Returns [(t, q)]."""
    out = []
    dtv = [dt_initial, dt_max]
    tend = t0
    dtout = (tfinal - t0) / float(nout)
    for _ in range(nout):
        tstart = tend
        tend = tstart + dtout
        r = claw3(p, q, aux, tstart, tend, dtv, cflv, nmax, backend=backend)
        if r.info != 0:
            raise RuntimeError(f"claw3 info={r.info}")
        q = r.q
        dtv = [r.dtv[4], dt_max]
        tend = r.tend
        out.append((tend, q))
    return out


def rpn3(ixyz, q_cells, aux_cells, mwaves=2, mbc=2):
    """The oracle's rpn3 over one slice: q_cells (n, meqn), aux_cells (n, maux)
    with cell index 1-mbc..n-mbc; entry k of the outputs is the interface
    between cells k-1 and k."""
    q = _prep(q_cells)
    a = _prep(aux_cells)
    n, meqn = q.shape
    maux = a.shape[1]
    mx = n - 2 * mbc
    wave = np.zeros((n, mwaves, meqn))
    s = np.zeros((n, mwaves))
    amdq = np.zeros((n, meqn))
    apdq = np.zeros((n, meqn))
    lib().orc_rpn3_vc_acoustics(ixyz, mx, meqn, mwaves, maux, mbc, mx, _dp(q), _dp(q), _dp(a),
                                _dp(a), _dp(wave), _dp(s), _dp(amdq), _dp(apdq))
    return wave, s, amdq, apdq


def rpt3(ixyz, icoor, imp, aux_slices, asdq, mwaves=2, mbc=2, impt=None):
    """rpt3 (or rptt3 when impt is given) over one slice.  aux_slices is
    (3, 3, n, maux): [aux1, aux2, aux3][kk][cell]; asdq (n, meqn)."""
    a = _prep(asdq)
    n, meqn = a.shape
    ax = _prep(aux_slices)
    maux = ax.shape[-1]
    mx = n - 2 * mbc
    bm = np.zeros((n, meqn))
    bp = np.zeros((n, meqn))
    a1, a2, a3 = (np.ascontiguousarray(ax[r]) for r in range(3))
    if impt is None:
        lib().orc_rpt3_vc_acoustics(ixyz, icoor, imp, mx, meqn, mwaves, maux, mbc, mx, None,
                                    None, _dp(a1), _dp(a2), _dp(a3), _dp(a), _dp(bm), _dp(bp))
    else:
        lib().orc_rptt3_vc_acoustics(ixyz, icoor, imp, impt, mx, meqn, mwaves, maux, mbc, mx,
                                     None, None, _dp(a1), _dp(a2), _dp(a3), _dp(a), _dp(bm),
                                     _dp(bp))
    return bm, bp


# ------------------------------------------------------------ driver checks

class OracleGrid3D:
    """The stepping interface of the device 3D grid (set_q / get_q / set_aux
    / claw3) on the C restatement, for checking driver logic on the CPU."""

    def __init__(self, mx, my, mz, maux=2, mbc=2, lower=(-1.0, -1.0, -1.0),
                 upper=(1.0, 1.0, 1.0)):
        self.mx, self.my, self.mz, self.maux, self.mbc = mx, my, mz, maux, mbc
        self.lower, self.upper = lower, upper
        self.q = self.aux = None

    def set_q(self, q):
        self.q = np.ascontiguousarray(q, dtype=np.float64).copy()

    def get_q(self):
        return self.q.copy()

    def set_aux(self, aux):
        self.aux = np.ascontiguousarray(aux, dtype=np.float64).copy()

    def claw3(self, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc, use_fwave=False):
        p = Problem3(self.mx, self.my, self.mz, maux=self.maux, mbc=self.mbc, method=method,
                     mthlim=mthlim, mthbc=mthbc, lower=self.lower, upper=self.upper,
                     use_fwave=use_fwave)
        r = claw3(p, self.q, self.aux, tstart, tend, dtv, cflv, nmax)
        self.q = r.q
        return r


def ref_claw3ez(rundir, q0, aux0):
    """Run the reference driver claw3ez (src/3d/claw3ez.f) in rundir with
    qinit/setaux returning q0/aux0 and the variable-coefficient acoustics
    solvers; it reads rundir/claw.data and writes fort.* there.  Needs
    oracle/_ref."""
    if not os.path.exists(os.path.join(rundir, "claw.data")):
        raise FileNotFoundError("claw.data missing (the reference driver would stop the process)")
    R = ref()
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    a0 = np.ascontiguousarray(aux0, dtype=np.float64)
    R.ref3_ez_set_init(_dp(q0), C.c_long(q0.size), _dp(a0), C.c_long(a0.size))
    cwd = os.getcwd()
    try:
        os.chdir(rundir)
        R.ref_claw3ez()
    finally:
        os.chdir(cwd)
        R.ref3_ez_set_init(None, C.c_long(0), None, C.c_long(0))
