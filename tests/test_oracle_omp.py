"""The OpenMP build of the oracle (the CPU baseline bench.py times) gives the
plain build's results bit for bit: its threads share the slices of a step by
bands (2D) or by the blocks of each pass (3D, as src/3d/step3.f90:132-869
does), and every cell still receives its terms in the serial order."""
import numpy as np
import pytest

from classic_amd import problems as P
from oracle import oracle as O
from oracle import oracle3 as O3


@pytest.mark.parametrize("threads", [2, 3, 8, 64])
@pytest.mark.parametrize("solver", ["euler", "acoustics"])
def test_step2_threads_bit_equal(threads, solver):
    mx, my = 37, 29
    if solver == "euler":
        q, dx, dy = P.euler_quadrants(mx, my, seed=5)
        p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                      method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 2, 3], mthbc=(1, 3, 1, 3))
        dt = 0.3 * dx
    else:
        q, dx, dy = P.acoustics_pulse(mx, my)
        p = O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                      method=(1, 2, 1, 0, 0, 0, 0), mthlim=[1, 4], mthbc=(3, 3, 2, 2))
        dt = 0.4 * dx / 2.0
    q = O.bc2(p, q)
    a, ca = O.step2(p, q, dx, dy, dt)
    b, cb = O.step2(p, q, dx, dy, dt, backend="omp", threads=threads)
    assert np.array_equal(a, b)
    assert ca == cb


def test_claw2_threads_bit_equal():
    mx, my = 48, 40
    q, dx, dy = P.euler_quadrants(mx, my, seed=2)
    p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                  method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4, 4], mthbc=(1, 1, 1, 1))
    # a first dt far too large: the run retakes, so the loop logic is covered too
    r1 = O.claw2(p, q, dx, dy, 0.0, 1e9, [5 * dx, 1e99], [1.0, 0.9], 12)
    r2 = O.claw2(p, q, dx, dy, 0.0, 1e9, [5 * dx, 1e99], [1.0, 0.9], 12, backend="omp",
                 threads=4)
    assert np.array_equal(r1.q, r2.q)
    assert np.array_equal(r1.dtv, r2.dtv) and np.array_equal(r1.cflv, r2.cflv)
    assert list(r1.nv) == list(r2.nv) and r1.info == r2.info


@pytest.mark.parametrize("threads", [2, 5])
def test_step3_threads_bit_equal(threads):
    n = 11
    q, aux, _ = P.acoustics3d_pulse(n, seed=4)
    p = O3.Problem3(n, n, n, mthbc=(3,) * 6, mthlim=(2, 2))
    q = O3.bc3(p, q)
    dt = 0.8 * p.geom.dx / 2.0
    a, ca = O3.step3(p, q, aux, dt)
    b, cb = O3.step3(p, q, aux, dt, backend="omp", threads=threads)
    assert np.array_equal(a, b)
    assert ca == cb
