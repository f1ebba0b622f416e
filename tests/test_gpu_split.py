"""Dimensionally split step (dimsp2 / step2ds, method(3) = -1 Godunov and -2
Strang) on the GPU against the oracle: one step and whole claw2 runs, both
solvers, bit for bit in exact arithmetic and to 1e-12 in fast arithmetic."""
import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from oracle import oracle as O
from tests.helpers import euler_problem, max_rel_err, smooth_acoustics

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _oprob(solver, mx, my, method, lim, mthbc):
    if solver == SOLVER_ACOUSTICS:
        return O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                         method=method, mthlim=lim, mthbc=mthbc)
    return O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4), method=method,
                     mthlim=lim, mthbc=mthbc)


@pytest.mark.parametrize("split", [-1, -2])
@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("order", [1, 2])
def test_dimsp2_single_step_euler(split, arith, order):
    mx, my = 131, 70
    q, dx, dy = euler_problem(mx, my, 5)
    method = (1, order, split, 0, 0, 0, 0)
    p = _oprob(SOLVER_EULER, mx, my, method, [4, 1, 2], (3, 1, 2, 2))
    q = O.bc2(p, q)
    dt = 0.2 * dx
    q_o, cfl_o = O.dimsp2(p, q, dx, dy, dt, [1.0, 0.9])
    with gpu.Grid2D(SOLVER_EULER, mx, my, [1.4, 1.0], arith=arith) as g:
        g.set_q(q)
        cfl_g = g.dimsp2(dx, dy, dt, method, [4, 1, 2], (1.0, 0.9))
        q_g = g.get_q()
    if arith == 0:
        assert cfl_g == cfl_o
        np.testing.assert_array_equal(q_g[2:-2, 2:-2], q_o[2:-2, 2:-2])
    else:
        assert abs(cfl_g - cfl_o) <= 1e-14 * cfl_o
        assert max_rel_err(q_g, q_o) <= 1e-14


@pytest.mark.parametrize("split", [-1, -2])
@pytest.mark.parametrize("arith", [0, 1])
def test_dimsp2_claw2_euler(split, arith):
    mx, my = 150, 131
    q, dx, dy = euler_problem(mx, my, 2)
    method, lim, mthbc = (1, 2, split, 0, 0, 0, 0), [4, 4, 4], (1, 1, 1, 1)
    p = _oprob(SOLVER_EULER, mx, my, method, lim, mthbc)
    r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 40)
    with gpu.Grid2D(SOLVER_EULER, mx, my, [1.4, 1.0], arith=arith) as g:
        g.set_q(q)
        out = g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 40, method, lim, mthbc)
        q_g = g.get_q()
    assert out["nv"][1] == r.nv[1] == 40
    if arith == 0:
        np.testing.assert_array_equal(q_g[2:-2, 2:-2], r.q[2:-2, 2:-2])
        np.testing.assert_array_equal(out["dtv"], r.dtv)
    else:
        assert max_rel_err(q_g, r.q) <= TOL


@pytest.mark.parametrize("split", [-1, -2])
def test_dimsp2_claw2_acoustics_walls_with_retake(split):
    """Acoustics with walls, started with a dt 10x too large: the first
    attempts are rejected and retaken (claw2.f:695-720)."""
    mx, my = 80, 64
    q = smooth_acoustics(mx, my, seed=2)
    method, lim, mthbc = (1, 2, split, 0, 0, 0, 0), [3, 3], (3, 3, 3, 3)
    p = _oprob(SOLVER_ACOUSTICS, mx, my, method, lim, mthbc)
    dx, dy = 1.0 / mx, 1.0 / my
    r = O.claw2(p, q, dx, dy, 0.0, 0.2, [10 * dx, 1e99], [1.0, 0.9], 5000)
    with gpu.Grid2D(SOLVER_ACOUSTICS, mx, my, [1.0, 4.0]) as g:
        g.set_q(q)
        out = g.claw2(dx, dy, 0.0, 0.2, [10 * dx, 1e99], [1.0, 0.9], 5000, method, lim, mthbc)
        q_g = g.get_q()
    assert out["info"] == r.info == 0
    assert out["nv"][1] == r.nv[1]
    np.testing.assert_array_equal(out["dtv"], r.dtv)
    np.testing.assert_array_equal(q_g[2:-2, 2:-2], r.q[2:-2, 2:-2])


def test_dimsp2_golden():
    """The reference build's own split-step vectors (tests/golden)."""
    import os
    for split in (1, 2):
        g = np.load(os.path.join(os.path.dirname(__file__), "golden",
                                 f"euler_dimsplit{split}_24x20.npz"))
        with gpu.Grid2D(SOLVER_EULER, 24, 20, [1.4, 1.0]) as gg:
            gg.set_q(g["q_in"])
            out = gg.claw2(float(g["dx"]), float(g["dy"]), 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 10,
                           (1, 2, -split, 0, 0, 0, 0), [4, 4, 4], (1, 1, 1, 1))
            q_g = gg.get_q()
        np.testing.assert_array_equal(q_g[2:-2, 2:-2], g["q_out"])
        np.testing.assert_array_equal(out["dtv"], g["dtv"])
