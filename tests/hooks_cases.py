"""claw2 runs with user routines (b4step2, src2) shared by the hook tests
and their golden generator: the test routines log (kind, t, dt) per call
(kind 1 = b4step2, 2 = src2) and src2 applies q(1) <- q(1) - (k dt) q(1)
on the interior cells."""
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


def python_hooks(k, mbc=2):
    """The test routines as Python callables for Grid2D.claw2 plus their log."""
    log = []

    def b4step2(q, aux, t, dt):
        log.append((1.0, t, dt))

    def src2(q, aux, t, dt):
        log.append((2.0, t, dt))
        kd = k * dt
        v = q[mbc:-mbc, mbc:-mbc, 0]
        q[mbc:-mbc, mbc:-mbc, 0] = v - kd * v

    return b4step2, src2, log
