"""claw1 / claw2 / claw3 runs with user routines (b4step, src) shared by the
hook tests and their golden generator: the test routines log (kind, t, dt)
per call (kind 1 = b4step, 2 = src) and src applies
q(1) <- q(1) - (k dt) q(1) on the interior cells."""
import numpy as np

from classic_amd import problems as P
from tests.helpers import smooth_acoustics

CASES = {
    # method(5) = 1 (source over dt after each accepted step), retakes
    "acoustics_src_godunov": dict(solver="acoustics", mx=30, my=24, method=(1, 2, 2, 0, 1, 0, 0),
                                  mthbc=(3, 3, 3, 3), dtv=(0.05, 1e99), tend=0.2, k=0.7, seed=11),
    # method(5) = 2 (Strang: dt/2 before and after), retakes restore the pre-source state
    "euler_src_strang": dict(solver="euler", mx=24, my=20, method=(1, 2, 2, 0, 2, 0, 0),
                             mthbc=(1, 1, 1, 1), dtv=(0.02, 1e99), tend=0.1, k=1.3, seed=12),
    # dimensional splitting (Strang sweeps) with the Godunov source step
    "euler_split_src": dict(solver="euler", mx=26, my=22, method=(1, 2, -2, 0, 1, 0, 0),
                            mthbc=(1, 1, 1, 1), dtv=(1e-3, 1e99), tend=0.08, k=0.5, seed=13),
}


def initial(case):
    c = CASES[case]
    if c["solver"] == "euler":
        q, dx, dy = P.euler_quadrants(c["mx"], c["my"], seed=c["seed"])
    else:
        q = smooth_acoustics(c["mx"], c["my"], seed=c["seed"])
        dx, dy = 1.0 / c["mx"], 1.0 / c["my"]
    return q, dx, dy


CASES1 = {
    # method(5) = 1; the first attempts are rejected (src1 runs on them too)
    "acoustics1d_src_godunov": dict(solver="acoustics", mx=80, maux=0, method=(1, 2, 0, 0, 1, 0, 0),
                                    mthlim=(4, 4), mthbc=(3, 3), dtv=(0.05, 1e99), tend=0.5,
                                    k=0.7, seed=21),
    # method(5) = 2 with variable coefficients, retakes restore the step's start
    "vc1d_src_strang": dict(solver="vc_acoustics", mx=64, maux=2, method=(1, 2, 0, 0, 2, 0, 0),
                            mthlim=(3, 3), mthbc=(1, 1), dtv=(0.04, 1e99), tend=0.3, k=1.3,
                            seed=22),
}

CASES3 = {
    # unsplit step, method(5) = 1 (q to 1e-12: the tile kernel's summation order)
    "acoustics3d_src_godunov": dict(mx=24, my=20, mz=16, method=(1, 2, 22, 0, 1, 0, 0),
                                    mthlim=(2, 2), mthbc=(3, 3, 3, 3, 3, 3), dtv=(0.2, 1e9),
                                    tend=0.3, k=0.7, seed=31),
    # split step, method(5) = 2 (bit for bit)
    "acoustics3d_src_strang_split": dict(mx=20, my=18, mz=14, method=(1, 2, -1, 0, 2, 0, 0),
                                         mthlim=(3, 3), mthbc=(1, 1, 3, 3, 1, 1), dtv=(0.3, 1e9),
                                         tend=0.3, k=1.3, seed=32),
}


def initial1(case):
    """(q, aux, dx) of a 1D case: random q and, for variable coefficients,
    random (Z, c) on [0, 1]."""
    c = CASES1[case]
    rng = np.random.default_rng(c["seed"])
    n = c["mx"] + 4
    q = rng.standard_normal((n, 2))
    aux = None
    if c["maux"]:
        aux = np.stack([rng.uniform(0.5, 3.0, n), rng.uniform(0.5, 2.0, n)], -1)
    return q, aux, 1.0 / c["mx"]


def initial3(case):
    """(q, aux) of a 3D case: a pressure pulse in the two-layer medium on
    [-1, 1]^3."""
    c = CASES3[case]
    mx, my, mz = c["mx"], c["my"], c["mz"]
    qc, _, _ = P.acoustics3d_pulse(mx, seed=c["seed"], width=0.3)
    q = np.ascontiguousarray(qc[: mz + 4, : my + 4])
    aux = P.layered_medium_aux(mx, my, mz, below=(1.0, 1.0), above=(6.0, 2.0))
    return q, aux


def python_hooks(k, mbc=2, dim=2):
    """The test routines as Python callables (b4step, src) for
    Grid1D.claw1 / Grid2D.claw2 / Grid3D.claw3 plus their log."""
    log = []
    inner = (slice(mbc, -mbc),) * dim + (0,)

    def b4step(q, aux, t, dt):
        log.append((1.0, t, dt))

    def src(q, aux, t, dt):
        log.append((2.0, t, dt))
        kd = k * dt
        v = q[inner]
        q[inner] = v - kd * v

    return b4step, src, log
