"""ctypes front end of the 1D CPU ORACLE -- test infrastructure only.

Loads ``oracle/_build/liboracle1d.so`` (``claw1d_oracle.c``) and, when built,
``oracle/_ref/libclawref1d.so`` (the reference Fortran from clawpack/classic
``src/1d`` linked with the same rp1 solvers).  Arrays use the reference layout
q(meqn, 1-mbc:mx+mbc): C-contiguous numpy arrays of shape (mx+2mbc, meqn).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle1d.so")
REF_SO = os.path.join(HERE, "_ref", "libclawref1d.so")

SOLVER_ACOUSTICS = 1
SOLVER_VC_ACOUSTICS = 3
SOLVER_ADVECTION = 4

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class Params1(C.Structure):
    _fields_ = [("rho", C.c_double), ("bulk", C.c_double), ("cc", C.c_double),
                ("zz", C.c_double), ("u", C.c_double)]


def params1(rho=1.0, bulk=4.0, u=1.0):
    cc = float(np.sqrt(bulk / rho))
    return Params1(rho, bulk, cc, cc * rho, u)


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


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
    return _ref


class Problem1:
    def __init__(self, mx, solver_id=SOLVER_ACOUSTICS, params=None, meqn=2, mwaves=2, maux=0,
                 mbc=2, method=(1, 2, 0, 0, 0, 0, 0), mthlim=(4, 4), mthbc=(1, 1),
                 use_fwave=False):
        self.mx, self.meqn, self.mwaves, self.maux, self.mbc = mx, meqn, mwaves, maux, mbc
        self.solver_id = solver_id
        self.params = params if params is not None else params1()
        m = list(method)
        m[6] = maux
        self.method = _arr_i(m)
        self.mthlim = _arr_i(mthlim)
        self.mthbc = _arr_i(mthbc)
        self.use_fwave = 1 if use_fwave else 0


def _prep(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _setp(p, backend):
    if backend == "oracle":
        lib().orc1_set_params(C.byref(p.params))
    else:
        ref().ref1_set_solver(p.solver_id, C.byref(p.params))


def step1(p: Problem1, q, dx, dt, aux=None, backend="oracle"):
    """One step (in place on a copy); returns (q, cfl)."""
    q = _prep(q).copy()
    a = _prep(aux) if aux is not None else None
    cfl = C.c_double(0.0)
    _setp(p, backend)
    if backend == "oracle":
        rc = lib().orc1_step1(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, _dp(q), _dp(a),
                              C.c_double(dx), C.c_double(dt), _ip(p.method), _ip(p.mthlim),
                              p.use_fwave, p.solver_id, C.byref(cfl))
    else:
        rc = ref().ref1_step1(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, _dp(q), _dp(a),
                              C.c_double(dx), C.c_double(dt), _ip(p.method), _ip(p.mthlim),
                              p.use_fwave, C.byref(cfl))
    if rc != 0:
        raise RuntimeError(f"step1 ({backend}) rc={rc}")
    return q, cfl.value


def bc1(p: Problem1, q, backend="oracle"):
    q = _prep(q).copy()
    f = lib().orc1_bc1 if backend == "oracle" else ref().ref1_bc1
    if f(p.meqn, p.mbc, p.mx, _dp(q), _ip(p.mthbc)) != 0:
        raise RuntimeError("bc1 failed")
    return q


class ClawResult:
    def __init__(self, q, info, tend, dtv, cflv, nv, dt_trace=None, cfl_trace=None):
        self.q, self.info, self.tend = q, info, tend
        self.dtv, self.cflv, self.nv = dtv, cflv, nv
        self.dt_trace, self.cfl_trace = dt_trace, cfl_trace


def claw1(p: Problem1, q, dx, tstart, tend, dtv, cflv, nmax, aux=None, backend="oracle",
          trace_cap=100000):
    q = _prep(q).copy()
    a = _prep(aux) if aux is not None else None
    te = C.c_double(tend)
    dtv_a = np.zeros(5)
    dtv_a[: len(dtv)] = dtv
    cflv_a = np.zeros(4)
    cflv_a[: len(cflv)] = cflv
    nv = np.array([nmax, 0], dtype=np.int32)
    _setp(p, backend)
    if backend == "oracle":
        dtt = np.zeros(trace_cap)
        cft = np.zeros(trace_cap)
        ntr = C.c_int(0)
        info = lib().orc1_claw1(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, _dp(q), _dp(a),
                                C.c_double(dx), C.c_double(tstart), C.byref(te), _dp(dtv_a),
                                _dp(cflv_a), _ip(nv), _ip(p.method), _ip(p.mthlim), _ip(p.mthbc),
                                p.use_fwave, p.solver_id, _dp(dtt), _dp(cft), trace_cap,
                                C.byref(ntr))
        return ClawResult(q, info, te.value, dtv_a, cflv_a, nv, dtt[: ntr.value], cft[: ntr.value])
    info = ref().ref1_claw1(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, _dp(q), _dp(a),
                            C.c_double(0.0), C.c_double(dx), C.c_double(tstart), C.byref(te),
                            _dp(dtv_a), _dp(cflv_a), _ip(nv), _ip(p.method), _ip(p.mthlim),
                            _ip(p.mthbc), p.use_fwave)
    return ClawResult(q, info, te.value, dtv_a, cflv_a, nv)


def frames(p: Problem1, q, dx, t0, tfinal, nout, dt_initial, cflv=(1.0, 0.9), dt_max=1e99,
           nmax=50000, aux=None, backend="oracle"):
    """claw1ez.f's output_style 1 loop (dtv(1) = dtv(5) between calls)."""
    out = []
    dtv = [dt_initial, dt_max]
    tend = t0
    dtout = (tfinal - t0) / float(nout)
    for _ in range(nout):
        tstart = tend
        r = claw1(p, q, dx, tstart, tstart + dtout, dtv, cflv, nmax, aux=aux, backend=backend)
        if r.info != 0:
            raise RuntimeError(f"claw1 info={r.info}")
        q = r.q
        dtv = [r.dtv[4], dt_max]
        tend = r.tend
        out.append((tend, q))
    return out


def rp1(solver_id, params, q_cells, aux_cells=None, mwaves=2, mbc=2):
    """The oracle's rp1 over one slice (q_cells (n, meqn)); entry k of the
    outputs is the interface between cells k-1 and k."""
    q = _prep(q_cells)
    n, meqn = q.shape
    a = _prep(aux_cells) if aux_cells is not None else None
    maux = a.shape[1] if a is not None else 0
    mx = n - 2 * mbc
    wave = np.zeros((n, mwaves, meqn))
    s = np.zeros((n, mwaves))
    amdq = np.zeros((n, meqn))
    apdq = np.zeros((n, meqn))
    L = lib()
    L.orc1_set_params(C.byref(params))
    fn = {SOLVER_ACOUSTICS: L.orc_rp1_acoustics, SOLVER_VC_ACOUSTICS: L.orc_rp1_acoustics_variable,
          SOLVER_ADVECTION: L.orc_rp1_advection}[solver_id]
    fn(mx, meqn, mwaves, maux, mbc, mx, _dp(q), _dp(q), _dp(a), _dp(a), _dp(wave), _dp(s),
       _dp(amdq), _dp(apdq))
    return wave, s, amdq, apdq


# ------------------------------------------------------------ driver checks

class OracleGrid1D:
    """The stepping interface of the device 1D grid (set_q / get_q / set_aux
    / claw1) on the C restatement, for checking driver logic on the CPU."""

    def __init__(self, solver_id, mx, params, mbc=2, maux=0, meqn=2, mwaves=2):
        self.solver_id, self.mx, self.mbc, self.maux = solver_id, mx, mbc, maux
        self.meqn, self.mwaves, self.params = meqn, mwaves, params
        self.q = self.aux = None

    def set_q(self, q):
        self.q = np.ascontiguousarray(q, dtype=np.float64).copy()

    def get_q(self):
        return self.q.copy()

    def set_aux(self, aux):
        self.aux = np.ascontiguousarray(aux, dtype=np.float64).copy()

    def claw1(self, dx, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc, use_fwave=False):
        p = Problem1(self.mx, self.solver_id, self.params, meqn=self.meqn, mwaves=self.mwaves,
                     maux=self.maux, mbc=self.mbc, method=method, mthlim=mthlim, mthbc=mthbc,
                     use_fwave=use_fwave)
        r = claw1(p, self.q, dx, tstart, tend, dtv, cflv, nmax, aux=self.aux)
        self.q = r.q
        return dict(info=r.info, tend=r.tend, dtv=r.dtv, cflv=r.cflv, nv=r.nv)


def ref_claw1ez(rundir, solver_id, params, q0, aux0=None):
    """Run the reference driver claw1ez (src/1d/claw1ez.f) in rundir with
    qinit/setaux returning q0/aux0; it reads rundir/claw.data and writes
    fort.* there.  Needs oracle/_ref."""
    if not os.path.exists(os.path.join(rundir, "claw.data")):
        raise FileNotFoundError("claw.data missing (the reference driver would stop the process)")
    R = ref()
    R.ref1_set_solver(solver_id, C.byref(params))
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    a0 = None if aux0 is None else np.ascontiguousarray(aux0, dtype=np.float64)
    R.ref1_ez_set_init(_dp(q0), C.c_long(q0.size), None if a0 is None else _dp(a0),
                       C.c_long(0 if a0 is None else a0.size))
    cwd = os.getcwd()
    try:
        os.chdir(rundir)
        R.ref_claw1ez()
    finally:
        os.chdir(cwd)
        R.ref1_ez_set_init(None, C.c_long(0), None, C.c_long(0))
