"""The claw*ez drivers: read claw.data, set up the grid, write frames.

``claw2ez`` follows src/2d/claw2ez.f:29-324, ``claw1ez`` src/1d/claw1ez.f
and ``claw3ez`` src/3d/claw3ez.f:51-360: read the run parameters, check the
periodic boundaries, compute dx = (upper - lower) / num_cells, call the
user's setprob / setaux / qinit (or read the restart frame), write frame 0,
then call the claw time loop once per output interval (output_style 1:
equal intervals accumulated as tend = tstart + dtout; 2: the listed times;
3: one step per call, a frame every output_step_interval steps), carrying
dtv(1) = dtv(5) from call to call, writing fort.qNNNN / fort.tNNNN (and
fort.aNNNN) and the fort.info record after each output, and stopping at the
first call that returns info != 0.

The stepping runs on the device (``Grid1D`` / ``Grid2D`` / ``Grid3D``).  The
user routines of the reference's plug-in contract become Python callables:

  qinit(meqn, mbc, mx, my, xlower, ylower, dx, dy, maux, aux) -> q
  setaux(mbc, mx, my, xlower, ylower, dx, dy, maux) -> aux
  setprob(rundir) -> None

(1D and 3D drop / add the corresponding coordinates), arrays in the
reference layout (numpy, C order, shape (my+2mbc, mx+2mbc, meqn) in 2D).
The Riemann solver, which the reference selects at link time, is chosen by
name: "acoustics" or "euler" (2D), "vc_advection" (2D, maux >= 2),
"vc_acoustics" (3D, maux = 2), "acoustics", "vc_acoustics" or "advection"
(1D), with its parameters in ``solver_params``.
"""
from __future__ import annotations

import os

import numpy as np

from . import output
from .rundata import read_claw_data


class RunResult:
    def __init__(self):
        self.frames = []   # (frame, t)
        self.info = 0
        self.t = 0.0
        self.records = []  # claw outputs per call: (tend, info, dtv, cflv, nv)


def _check_periodic(mthbc):
    for a in range(0, len(mthbc), 2):
        if (mthbc[a] == 2) != (mthbc[a + 1] == 2):
            raise ValueError("periodic boundary conditions require both sides of an axis set to 2")


def _out_times(c):
    """tend of call n (1-based) given tstart, per output style."""
    if c.outstyle == 1:
        dtout = (c.tfinal - c.t0) / float(c.nout)
        return lambda n, tstart: tstart + dtout
    if c.outstyle == 2:
        return lambda n, tstart: c.tout[n - 1]
    return lambda n, tstart: tstart - 1.0  # single-step mode


class _Stepper:
    """The device grid behind one driver run (dimension-specific calls)."""

    def __init__(self, ndim, c, solver, solver_params, arith, grid=None):
        self.ndim = ndim
        self.c = c
        n = c.num_cells
        self.delta = [(c.upper[a] - c.lower[a]) / float(n[a]) for a in range(ndim)]
        if grid is not None:
            self.g = grid
            return
        if ndim == 2:
            from . import gpu
            from ._lib import SOLVER_ACOUSTICS, SOLVER_EULER, SOLVER_VC_ADVECTION
            sid = {"acoustics": SOLVER_ACOUSTICS, "euler": SOLVER_EULER,
                   "vc_advection": SOLVER_VC_ADVECTION}[solver]
            params = list(solver_params) if solver_params is not None else (
                [] if sid == SOLVER_VC_ADVECTION else gpu.solver_params(sid))
            self.g = gpu.Grid2D(sid, n[0], n[1], params if params else [0.0, 0.0], mbc=c.mbc,
                                maux=c.maux, arith=arith)
        elif ndim == 1:
            from . import gpu1
            from ._lib import SOLVER_ACOUSTICS, SOLVER_ADVECTION, SOLVER_VC_ACOUSTICS
            sid = {"acoustics": SOLVER_ACOUSTICS, "vc_acoustics": SOLVER_VC_ACOUSTICS,
                   "advection": SOLVER_ADVECTION}[solver]
            self.g = gpu1.Grid1D(sid, n[0], tuple(solver_params or ()), maux=c.maux, mbc=c.mbc)
        else:
            from . import gpu3
            if solver not in ("vc_acoustics", None):
                raise ValueError("the 3D path runs the variable-coefficient acoustics solver")
            self.g = gpu3.Grid3D(n[0], n[1], n[2], maux=c.maux, mbc=c.mbc, lower=c.lower,
                                 upper=c.upper)
            # claw3ez.f:216-218 computes dx, dy, dz as (upper - lower) / n
            self.g.geom[3:] = self.delta

    def close(self):
        if hasattr(self.g, "close"):
            self.g.close()

    def claw(self, tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc, use_fwave):
        """One call of the claw time loop; returns (tend, info, dtv, cflv, nv)."""
        kw = dict(getattr(self, "hooks", {}))
        if self.ndim == 2:
            if kw:
                kw.update(xlower=self.c.lower[0], ylower=self.c.lower[1])
            r = self.g.claw2(self.delta[0], self.delta[1], tstart, tend, dtv, cflv, nmax, method,
                             mthlim, mthbc, use_fwave, **kw)
            return r["tend"], r["info"], r["dtv"], r["cflv"], r["nv"]
        if self.ndim == 1:
            if kw:
                kw.update(xlower=self.c.lower[0])
            r = self.g.claw1(self.delta[0], tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc,
                             use_fwave, **kw)
            return r["tend"], r["info"], r["dtv"], r["cflv"], r["nv"]
        r = self.g.claw3(tstart, tend, dtv, cflv, nmax, method, mthlim, mthbc, use_fwave, **kw)
        return r.tend, r.info, r.dtv, r.cflv, r.nv


ROUTINES = {1: ("bc1", "b4step1", "src1"), 2: ("bc2", "b4step2", "src2"),
            3: ("bc3", "b4step3", "src3")}


def run(rundir=".", outdir=None, qinit=None, setaux=None, setprob=None, solver="acoustics",
        solver_params=None, arith=0, grid=None, verbose=False, **routines):
    """Run the claw*ez driver for the claw.data in rundir; frames go to
    outdir (default: rundir).  Returns a RunResult.  ``grid`` may supply an
    already constructed stepping object with the Grid*D interface.  The
    keyword arguments bc<d> / b4step<d> / src<d> (d = the dimension) are the
    user routines (see gpu.make_hooks, gpu1.make_hooks1, gpu3.make_hooks3);
    bc<d> is required when a boundary code is 0 (user boundary), except for
    the 3D piston wall (mthbc(6) = 0)."""
    c = read_claw_data(os.path.join(rundir, "claw.data"))
    nd = c.ndim
    known = {k for names in ROUTINES.values() for k in names}
    unknown = sorted(set(routines) - known)
    if unknown:
        raise TypeError(f"unexpected keyword arguments {unknown}")
    hooks = {k: v for k, v in routines.items() if v is not None}
    if any(k not in ROUTINES[nd] for k in hooks):
        raise ValueError(f"the user routines of a {nd}D run are {', '.join(ROUTINES[nd])}")
    bc = ROUTINES[nd][0]
    if 0 in c.mthbc and bc not in hooks and not (nd == 3 and c.mthbc == [1, 1, 1, 1, 1, 0]):
        raise ValueError(f"a boundary code 0 (user boundary) needs the {bc} routine")
    if nd == 1:
        c["restart"] = False  # claw1ez reads no restart record
    outdir = outdir or rundir
    os.makedirs(outdir, exist_ok=True)
    _check_periodic(c.mthbc)
    if c.mbc < 2:
        raise ValueError("num_ghost must be >= 2 (claw info = 1)")
    if qinit is None and not c.restart:
        raise ValueError("qinit is required unless the run restarts from a frame")
    st = _Stepper(nd, c, solver, solver_params, arith, grid)
    st.hooks = hooks
    res = RunResult()
    info_path = os.path.join(outdir, "fort.info")
    open(info_path, "w").close()  # claw2ez.f:35 opens fort.info anew
    try:
        if setprob is not None:
            setprob(rundir)
        lower, delta, n = list(c.lower), st.delta, c.num_cells
        geo = []
        if nd == 1:
            geo = [n[0], lower[0], delta[0]]
        elif nd == 2:
            geo = [n[0], n[1], lower[0], lower[1], delta[0], delta[1]]
        else:
            geo = [n[0], n[1], n[2], lower[0], lower[1], lower[2], delta[0], delta[1], delta[2]]
        aux = None
        if c.maux > 0:
            if setaux is None:
                raise ValueError("num_aux > 0 needs setaux")
            aux = np.ascontiguousarray(setaux(c.mbc, *geo, c.maux), dtype=np.float64)
        if c.restart:
            t0, q = output.read_frame(outdir, c.restart_frame, nd, c.meqn, c.mbc)
            iframe = c.restart_frame
        else:
            q = np.ascontiguousarray(qinit(c.meqn, c.mbc, *geo, c.maux, aux), dtype=np.float64)
            t0 = c.t0
            iframe = 0
            # out2 flushes tiny values of q (and aux) in place before the run starts
            output.write_frame(outdir, 0, t0, q, c.mbc, lower, delta, aux=aux,
                               write_aux=c.outaux_init_only or c.outaux_always)
            res.frames.append((0, t0))
        if aux is not None:
            st.g.set_aux(aux)
        st.g.set_q(q)
        dtv = list(c.dtv[:2])
        cflv = list(c.cflv[:2])
        tend_of = _out_times(c)
        tend = t0
        n0 = iframe * c.nstepout + 1
        for ncall in range(n0, c.nout + 1):
            tstart = tend
            tend = tend_of(ncall, tstart)
            te, info, dtv_o, cflv_o, nv = st.claw(tstart, tend, dtv, cflv, c.nv[0], c.method,
                                                  c.mthlim, c.mthbc, c.use_fwaves)
            res.records.append((te, info, list(dtv_o), list(cflv_o), list(nv)))
            tend = te
            res.t = te
            if info != 0:
                res.info = info
                if verbose:
                    print(f"claw{nd}ez aborting: error return from claw{nd} with info = {info}")
                break
            dtv[0] = float(dtv_o[4])  # the final dt starts the next call (claw2ez.f:283)
            if ncall % c.nstepout == 0:
                frame = ncall // c.nstepout
                qf = st.g.get_q()
                flushed = qf.copy()
                aux_before = aux.copy() if (aux is not None and c.outaux_always) else None
                output.write_frame(outdir, frame, te, flushed, c.mbc, lower, delta, aux=aux,
                                   write_aux=c.outaux_always)
                if not np.array_equal(flushed, qf):
                    st.g.set_q(flushed)  # the reference continues from the flushed array
                if aux_before is not None and not np.array_equal(aux, aux_before):
                    st.g.set_aux(aux)
                with open(info_path, "a") as f:
                    f.write(output.info_record(te, info, dtv_o, cflv_o, nv))
                res.frames.append((frame, te))
                if verbose:
                    print(f"CLAW{nd}EZ: Frame {frame:4d} output files done at time t = {te:12.4e}")
    finally:
        if grid is None:
            st.close()
    return res


def claw1ez(rundir=".", **kw):
    return run(rundir, **kw)


def claw2ez(rundir=".", **kw):
    return run(rundir, **kw)


def claw3ez(rundir=".", **kw):
    kw.setdefault("solver", "vc_acoustics")
    return run(rundir, **kw)
