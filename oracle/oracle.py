"""ctypes front end of the CPU ORACLE -- test infrastructure only.

Loads ``oracle/_build/liboracle2d.so`` (the plain-C restatement in
``claw2d_oracle.c``) and, when it was built, ``oracle/_ref/libclawref2d.so``
(the reference Fortran from clawpack/classic ``src/2d`` linked with the same
Riemann solvers).  Only tests, the smoke check and the CPU-baseline leg of
bench.py import this module; the product package never does.

Arrays use the reference layout q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc) in
Fortran order, i.e. a C-contiguous numpy array of shape (my+2mbc, mx+2mbc,
meqn).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle2d.so")
REF_SO = os.path.join(HERE, "_ref", "libclawref2d.so")
ORACLE_OMP_SO = os.path.join(HERE, "_build", "liboracle2d_omp.so")

SOLVER_ACOUSTICS = 1
SOLVER_EULER = 2
SOLVER_VC_ADVECTION = 5


class SolverParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("bulk", C.c_double), ("cc", C.c_double),
                ("zz", C.c_double), ("gamma", C.c_double), ("efix", C.c_int)]


def solver_params(rho=1.0, bulk=4.0, gamma=1.4, efix=True):
    cc = float(np.sqrt(bulk / rho))
    return SolverParams(rho, bulk, cc, rho * cc, gamma, 1 if efix else 0)


def build(ref: bool = False) -> None:
    """Compile the oracle (and the reference build when ``ref``)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_set_solver_params.argtypes = [C.POINTER(SolverParams)]
    return _lib


_omp = None


def lib_omp(threads: int):
    """The OpenMP build of the same restatement (-O3, threaded unsplit steps
    that update every cell in the serial order, so results are the plain
    build's bit for bit), set to `threads` threads; backend "omp" below."""
    global _omp
    if _omp is None:
        if not os.path.exists(ORACLE_OMP_SO):
            build()
        _omp = C.CDLL(ORACLE_OMP_SO)
        _omp.orc_set_solver_params.argtypes = [C.POINTER(SolverParams)]
    _omp.orc_set_threads(int(threads))
    return _omp


def _cpu(backend, threads):
    return lib() if backend == "oracle" else lib_omp(threads)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
    return _ref


def _arr_i(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _aux_ptr(aux):
    return _dp(aux) if aux is not None else None


class Problem:
    """Sizes, method flags and solver of one 2D run (reference argument set)."""

    def __init__(self, meqn, mwaves, mx, my, solver_id, params, maux=0, mbc=2,
                 method=(1, 2, 2, 0, 0, 0, 0), mthlim=None, mthbc=(1, 1, 1, 1),
                 use_fwave=False):
        self.meqn, self.mwaves, self.mx, self.my = meqn, mwaves, mx, my
        self.maux, self.mbc = maux, mbc
        self.solver_id, self.params = solver_id, params
        m = list(method)
        m[6] = maux
        self.method = _arr_i(m)
        self.mthlim = _arr_i(mthlim if mthlim is not None else [4] * mwaves)
        self.mthbc = _arr_i(mthbc)
        self.use_fwave = 1 if use_fwave else 0

    def shape(self):
        return (self.my + 2 * self.mbc, self.mx + 2 * self.mbc, self.meqn)


def _prep(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return q


def step2(p: Problem, qold, dx, dy, dt, aux=None, backend="oracle", threads=1):
    """One unsplit step; returns (qnew, cfl).  qnew starts as a copy of qold.
    backend: "oracle" (the C restatement), "omp" (its OpenMP build on
    `threads` threads) or "ref" (the reference build)."""
    qold = _prep(qold)
    qnew = qold.copy()
    cfl = C.c_double(0.0)
    if backend in ("oracle", "omp"):
        L = _cpu(backend, threads)
        L.orc_set_solver_params(C.byref(p.params))
        rc = L.orc_step2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(qold), _dp(qnew),
                         _aux_ptr(aux), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                         _ip(p.method), _ip(p.mthlim), p.use_fwave, p.solver_id, C.byref(cfl))
    else:
        R = ref()
        R.ref_set_solver(p.solver_id, C.byref(p.params))
        rc = R.ref_step2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(qold), _dp(qnew),
                         _aux_ptr(aux), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                         _ip(p.method), _ip(p.mthlim), p.use_fwave, C.byref(cfl))
    if rc != 0:
        raise RuntimeError(f"step2 ({backend}) failed rc={rc}")
    return qnew, cfl.value


def dimsp2(p: Problem, qold, dx, dy, dt, cflv, aux=None, backend="oracle"):
    qold = _prep(qold)
    qnew = qold.copy()
    cfl = C.c_double(0.0)
    cv = np.ascontiguousarray(np.asarray(cflv, dtype=np.float64))
    if backend == "oracle":
        L = lib()
        L.orc_set_solver_params(C.byref(p.params))
        rc = L.orc_dimsp2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(qold), _dp(qnew),
                          _aux_ptr(aux), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                          _ip(p.method), _ip(p.mthlim), _dp(cv), p.use_fwave, p.solver_id,
                          C.byref(cfl))
    else:
        R = ref()
        R.ref_set_solver(p.solver_id, C.byref(p.params))
        rc = R.ref_dimsp2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(qold), _dp(qnew),
                          _aux_ptr(aux), C.c_double(dx), C.c_double(dy), C.c_double(dt),
                          _ip(p.method), _ip(p.mthlim), _dp(cv), p.use_fwave, C.byref(cfl))
    if rc != 0:
        raise RuntimeError(f"dimsp2 ({backend}) failed rc={rc}")
    return qnew, cfl.value


def bc2(p: Problem, q, backend="oracle"):
    q = _prep(q).copy()
    if backend == "oracle":
        rc = lib().orc_bc2(p.meqn, p.mbc, p.mx, p.my, _dp(q), _ip(p.mthbc))
    else:
        rc = ref().ref_bc2(p.meqn, p.mbc, p.mx, p.my, _dp(q), _ip(p.mthbc))
    if rc != 0:
        raise RuntimeError("bc2 failed")
    return q


class ClawResult:
    def __init__(self, q, info, tend, dtv, cflv, nv, dt_trace=None, cfl_trace=None):
        self.q, self.info, self.tend = q, info, tend
        self.dtv, self.cflv, self.nv = dtv, cflv, nv
        self.dt_trace, self.cfl_trace = dt_trace, cfl_trace


def claw2(p: Problem, q, dx, dy, tstart, tend, dtv, cflv, nmax, aux=None,
          xlower=0.0, ylower=0.0, backend="oracle", trace_cap=100000, threads=1):
    """The claw2 time loop (src/2d/claw2.f:494-784).  Returns ClawResult.
    backend as for step2."""
    q = _prep(q).copy()
    te = C.c_double(tend)
    dtv_a = np.zeros(5)
    dtv_a[: len(dtv)] = dtv
    cflv_a = np.zeros(4)
    cflv_a[: len(cflv)] = cflv
    nv = np.array([nmax, 0], dtype=np.int32)
    if backend in ("oracle", "omp"):
        L = _cpu(backend, threads)
        L.orc_set_solver_params(C.byref(p.params))
        dtt = np.zeros(trace_cap)
        cft = np.zeros(trace_cap)
        ntr = C.c_int(0)
        info = L.orc_claw2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(q), _aux_ptr(aux),
                           C.c_double(dx), C.c_double(dy), C.c_double(tstart), C.byref(te),
                           _dp(dtv_a), _dp(cflv_a), _ip(nv), _ip(p.method), _ip(p.mthlim),
                           _ip(p.mthbc), p.use_fwave, p.solver_id, _dp(dtt), _dp(cft),
                           trace_cap, C.byref(ntr))
        return ClawResult(q, info, te.value, dtv_a, cflv_a, nv, dtt[: ntr.value],
                          cft[: ntr.value])
    R = ref()
    R.ref_set_solver(p.solver_id, C.byref(p.params))
    info = R.ref_claw2(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, _dp(q), _aux_ptr(aux),
                       C.c_double(xlower), C.c_double(ylower), C.c_double(dx), C.c_double(dy),
                       C.c_double(tstart), C.byref(te), _dp(dtv_a), _dp(cflv_a), _ip(nv),
                       _ip(p.method), _ip(p.mthlim), _ip(p.mthbc), p.use_fwave)
    return ClawResult(q, info, te.value, dtv_a, cflv_a, nv)


def rpn2(solver_id, params, ixy, ql_cells, mwaves, mbc=2):
    """Run the oracle's rpn2 over a slice of cells (shape (n, meqn), cell
    index 1-mbc..n-mbc); returns wave (n, mwaves, meqn), s (n, mwaves), amdq,
    apdq (n, meqn); entry k describes the interface between cells k-1, k."""
    q = _prep(ql_cells)
    n, meqn = q.shape
    mx = n - 2 * mbc
    wave = np.zeros((n, mwaves, meqn))
    s = np.zeros((n, mwaves))
    amdq = np.zeros((n, meqn))
    apdq = np.zeros((n, meqn))
    L = lib()
    L.orc_set_solver_params(C.byref(params))
    fn = L.orc_rpn2_acoustics if solver_id == SOLVER_ACOUSTICS else L.orc_rpn2_euler
    fn(ixy, mx, meqn, mwaves, 0, mbc, mx, _dp(q), _dp(q), None, None, _dp(wave), _dp(s),
       _dp(amdq), _dp(apdq))
    return wave, s, amdq, apdq


def rpt2(solver_id, params, ixy, imp, ql_cells, asdq, mwaves, mbc=2):
    q = _prep(ql_cells)
    a = _prep(asdq)
    n, meqn = q.shape
    mx = n - 2 * mbc
    bm = np.zeros((n, meqn))
    bp = np.zeros((n, meqn))
    L = lib()
    L.orc_set_solver_params(C.byref(params))
    fn = L.orc_rpt2_acoustics if solver_id == SOLVER_ACOUSTICS else L.orc_rpt2_euler
    fn(ixy, imp, mx, meqn, mwaves, 0, mbc, mx, _dp(q), _dp(q), None, None, None, _dp(a),
       _dp(bm), _dp(bp))
    return bm, bp


def limiter(wave, s, mthlim, mbc=2):
    """Apply the oracle limiter to wave (n, mwaves, meqn) in place-copy."""
    w = _prep(wave).copy()
    s = _prep(s)
    n, mwaves, meqn = w.shape
    mx = n - 2 * mbc
    rc = lib().orc_limiter(mx, meqn, mwaves, mbc, mx, _dp(w), _dp(s), _ip(_arr_i(mthlim)))
    if rc != 0:
        raise ValueError("invalid limiter")
    return w


def annulus_aux(mx, my, mbc=2, xlower=0.2, xupper=1.0, ylower=0.0, yupper=2.0 * np.pi):
    """aux(3, 1-mbc:mx+mbc, 1-mbc:my+mbc) of examples/advection_2d_annulus
    (edge velocities from the stream function, capacity = cell area ratio),
    computed by the oracle's C restatement of its setaux.f."""
    dxc = (xupper - xlower) / mx
    dyc = (yupper - ylower) / my
    aux = np.zeros((my + 2 * mbc, mx + 2 * mbc, 3))
    lib().orc_annulus_aux(mbc, mx, my, C.c_double(xlower), C.c_double(ylower), C.c_double(dxc),
                          C.c_double(dyc), _dp(aux))
    return aux, dxc, dyc


# ------------------------------------------------------------ driver checks

SOLVER_BY_NAME = {"acoustics": SOLVER_ACOUSTICS, "euler": SOLVER_EULER}
SHAPE_BY_ID = {SOLVER_ACOUSTICS: (3, 2), SOLVER_EULER: (4, 3)}


class OracleGrid2D:
    """The stepping interface of the device grid (set_q / get_q / set_aux /
    claw2) on the C restatement, for checking driver logic on the CPU."""

    def __init__(self, solver_id, mx, my, params, mbc=2, maux=0):
        self.solver_id, self.mx, self.my, self.mbc, self.maux = solver_id, mx, my, mbc, maux
        self.meqn, self.mwaves = SHAPE_BY_ID[solver_id]
        self.params = params
        self.q = None
        self.aux = None

    def set_q(self, q):
        self.q = np.ascontiguousarray(q, dtype=np.float64).copy()

    def get_q(self):
        return self.q.copy()

    def set_aux(self, aux):
        self.aux = np.ascontiguousarray(aux, dtype=np.float64).copy()

    def claw2(self, dx, dy, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc,
              use_fwave=False):
        p = Problem(self.meqn, self.mwaves, self.mx, self.my, self.solver_id, self.params,
                    maux=self.maux, mbc=self.mbc, method=method, mthlim=mthlim, mthbc=mthbc,
                    use_fwave=use_fwave)
        r = claw2(p, self.q, dx, dy, tstart, tend, dtv, cflv, nmax, aux=self.aux)
        self.q = r.q
        return dict(info=r.info, tend=r.tend, dtv=r.dtv, cflv=r.cflv, nv=r.nv)


def ref_claw2ez(rundir, solver_id, params, q0, aux0=None):
    """Run the reference driver claw2ez (src/2d/claw2ez.f) in rundir with
    qinit/setaux returning q0/aux0 and the given solver; it reads
    rundir/claw.data and writes fort.* there.  Needs oracle/_ref."""
    if not os.path.exists(os.path.join(rundir, "claw.data")):
        raise FileNotFoundError("claw.data missing (the reference driver would stop the process)")
    R = ref()
    R.ref_set_solver(solver_id, C.byref(params))
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    a0 = None if aux0 is None else np.ascontiguousarray(aux0, dtype=np.float64)
    R.ref_ez_set_init(_dp(q0), C.c_long(q0.size), _aux_ptr(a0), C.c_long(0 if a0 is None else a0.size))
    cwd = os.getcwd()
    try:
        os.chdir(rundir)
        R.ref_claw2ez()
    finally:
        os.chdir(cwd)
        R.ref_ez_set_init(None, C.c_long(0), None, C.c_long(0))


def ref_out2(outdir, q, t, frame, mbc, lower, delta, aux=None, outaux=False):
    """Write fort.q/t(/a)NNNN with the reference's out2 (src/2d/out2.f);
    note that it flushes |q| < 1e-99 to zero in q, as the reference does."""
    R = ref()
    q = np.ascontiguousarray(q, dtype=np.float64)
    my2, mx2, meqn = q.shape
    maux = 0 if aux is None else aux.shape[-1]
    a = aux if aux is not None else np.zeros(1)
    i = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
    d = lambda v: C.byref(C.c_double(float(v)))  # noqa: E731
    cwd = os.getcwd()
    try:
        os.chdir(outdir)
        R.out2_(i(meqn), i(mbc), i(mx2 - 2 * mbc), i(my2 - 2 * mbc), d(lower[0]), d(lower[1]),
                d(delta[0]), d(delta[1]), _dp(q), d(t), i(frame), _dp(a), i(maux),
                i(1 if outaux else 0))
    finally:
        os.chdir(cwd)
    return q
