"""Device Riemann solvers (clawgpu_solver_probe) against the oracle's rpn2 /
rpt2, interface by interface: bit for bit in exact arithmetic, to a few ulps
in fast arithmetic.  The data include every branch of the Harten-Hyman
entropy fix: ordinary, transonic (Toro test 1 rarefaction) and fully
supersonic interfaces."""
import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu

G = 1.4


def cons(rho, u, v, p):
    return np.stack([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)


def euler_rows():
    rng = np.random.default_rng(11)
    n = 400
    rows = [cons(rng.uniform(0.1, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                 rng.uniform(0.1, 2, n))]
    # supersonic flow in both directions and a transonic expansion
    x = np.linspace(0, 1, n)
    rows.append(cons(1 + 0.5 * x, 2.5 + 0 * x, 0.3 + 0 * x, 1 + 0.2 * x))
    rows.append(cons(1 + 0.5 * x, -2.5 + 0 * x, -0.3 + 0 * x, 1 + 0.2 * x))
    rows.append(cons(1 - 0.6 * x, 0.2 + 1.6 * x, 0 * x, 1 - 0.8 * x))
    rows.append(cons(1 - 0.6 * x, -0.2 - 1.6 * x, 0 * x, 1 - 0.8 * x))
    # Toro test 1 after a few steps of the reference arithmetic
    from tests.test_gpu_step2 import _oracle_problem, _toro1
    mx, my = 120, 6
    q, dx, dy = _toro1(mx, my)
    p = _oracle_problem(SOLVER_EULER, mx, my, (1, 2, 2, 0, 0, 0, 0), [4, 4, 4])
    for nst in (1, 2, 3, 6):
        r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], nst)
        rows.append(O.bc2(p, r.q)[4])
    return rows


def _compare(solver, params_o, params_g, rows, ixy, arith):
    with gpu.Grid2D(solver, 8, 8, params_g, arith=arith) as g:
        for row in rows:
            n, meqn = row.shape
            rng = np.random.default_rng(n)
            asdq = rng.standard_normal((n, meqn))
            mwaves = g.mwaves
            w_o, s_o, am_o, ap_o = O.rpn2(solver, params_o, ixy, row, mwaves)
            bm_o, bp_o = O.rpt2(solver, params_o, ixy, 1, row, asdq, mwaves)
            w_g, s_g, am_g, ap_g, bm_g, bp_g = g.solver_probe(ixy, row[:-1], row[1:], asdq[1:])
            pairs = [(w_g, w_o[1:]), (s_g, s_o[1:]), (am_g, am_o[1:]), (ap_g, ap_o[1:]),
                     (bm_g, bm_o[1:]), (bp_g, bp_o[1:])]
            for a, b in pairs:
                if arith == 0:
                    np.testing.assert_array_equal(a, b)
                else:
                    scale = np.abs(b).max() + 1e-300
                    assert np.abs(a - b).max() <= 1e-13 * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("ixy", [1, 2])
def test_euler_solver(arith, ixy):
    rows = euler_rows()
    if ixy == 2:
        rows = [r[:, [0, 2, 1, 3]] for r in rows]  # same physics along y
    _compare(SOLVER_EULER, O.solver_params(gamma=G, efix=True), [G, 1.0], rows, ixy, arith)


@pytest.mark.parametrize("arith", [0, 1])
def test_euler_solver_no_efix(arith):
    _compare(SOLVER_EULER, O.solver_params(gamma=G, efix=False), [G, 0.0], euler_rows()[:3], 1,
             arith)


@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("ixy", [1, 2])
def test_acoustics_solver(arith, ixy):
    rng = np.random.default_rng(5)
    rows = [rng.standard_normal((300, 3))]
    _compare(SOLVER_ACOUSTICS, O.solver_params(2.0, 5.0), [2.0, 5.0], rows, ixy, arith)
