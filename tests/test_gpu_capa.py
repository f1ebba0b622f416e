"""General 2D step on the GPU (aux arrays and capacity function,
variable-coefficient advection) against the oracle, bit for bit: single steps
for every method(2)/method(3)/limiter combination with and without capacity,
a claw2 run with retakes, and the annulus example with its regression sums."""
import os

import numpy as np
import pytest

from classic_amd import gpu, problems as P
from classic_amd._lib import SOLVER_VC_ADVECTION
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _random(mx, my, seed, capa):
    rng = np.random.default_rng(seed)
    aux = np.stack([rng.uniform(-1, 1, (my + 4, mx + 4)), rng.uniform(-1, 1, (my + 4, mx + 4)),
                    rng.uniform(0.5, 1.5, (my + 4, mx + 4))], -1)
    q = rng.standard_normal((my + 4, mx + 4, 1))
    return q, aux


@pytest.mark.parametrize("capa", [0, 3])
@pytest.mark.parametrize("order,trans,lim,fw", [(2, 2, 4, 0), (2, 1, 3, 0), (2, 0, 1, 0),
                                                (1, 2, 0, 0), (1, 1, 0, 0), (2, 2, 2, 1)])
def test_step2_general_bitwise(capa, order, trans, lim, fw):
    mx, my = 131, 70
    q, aux = _random(mx, my, 10 * order + trans + lim + capa, capa)
    method = (1, order, trans, 0, 0, capa, 3)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3, method=method,
                  mthlim=[lim], mthbc=(1, 1, 2, 2), use_fwave=bool(fw))
    q = O.bc2(p, q)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / my
    ref, cfl_ref = O.step2(p, q, dx, dy, dt, aux=aux)
    with gpu.Grid2D(SOLVER_VC_ADVECTION, mx, my, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(q)
        cfl = g.step2(dx, dy, dt, method, [lim], use_fwave=bool(fw))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2], ref[2:-2, 2:-2])


def test_claw2_general_with_retakes_bitwise():
    mx, my = 64, 48
    q, aux = _random(mx, my, 3, 3)
    method = (1, 2, 2, 0, 0, 3, 3)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3, method=method,
                  mthlim=[4], mthbc=(1, 1, 2, 2))
    r = O.claw2(p, q, 1.0 / mx, 1.0 / my, 0.0, 0.2, [0.1, 1e9], [1.0, 0.9], 1000, aux=aux)
    with gpu.Grid2D(SOLVER_VC_ADVECTION, mx, my, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(q)
        rg = g.claw2(1.0 / mx, 1.0 / my, 0.0, 0.2, [0.1, 1e9], [1.0, 0.9], 1000, method, [4],
                     (1, 1, 2, 2))
        out = g.get_q()
    assert rg["info"] == r.info == 0 and rg["tend"] == r.tend
    np.testing.assert_array_equal(rg["dtv"], r.dtv)
    np.testing.assert_array_equal(rg["cflv"], r.cflv)
    np.testing.assert_array_equal(rg["nv"], r.nv)
    assert len(r.dt_trace) > r.nv[1]
    np.testing.assert_array_equal(out[2:-2, 2:-2], r.q[2:-2, 2:-2])


def test_annulus_example_bitwise_and_regression_sums():
    gold = np.load(os.path.join(GOLD, "annulus_40x120.npz"))
    aux, dxc, dyc = O.annulus_aux(40, 120)
    q = P.annulus_hump(40, 120)
    method = (1, 2, 2, 0, 0, 3, 3)
    with gpu.Grid2D(SOLVER_VC_ADVECTION, 40, 120, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(q)
        dtv, t = [0.1, 1e9], 0.0
        for k in range(2):
            r = g.claw2(dxc, dyc, t, t + 0.25, dtv, [1.0, 0.9], 50000, method, [3], (1, 1, 2, 2))
            assert r["info"] == 0
            dtv, t = [r["dtv"][4], 1e9], r["tend"]
            qk = g.get_q()
            assert t == gold["t"][k]
            np.testing.assert_array_equal(qk[2:-2, 2:-2], gold["q"][k])
            assert abs(qk[2:-2, 2:-2, 0].sum() - gold["regression_sums"][k]) < 1e-10


@pytest.mark.parametrize("m3", [-1, -2])
@pytest.mark.parametrize("capa", [0, 3])
@pytest.mark.parametrize("order,lim,fw", [(2, 4, 0), (2, 3, 0), (1, 0, 0), (2, 2, 1), (2, 1, 0)])
def test_dimsp2_general_bitwise(m3, capa, order, lim, fw):
    """Split step with aux arrays and the capacity function (step2ds with
    index_capa > 0) against the oracle, bit for bit, Courant number
    included; grid with ragged 64-lane rows and 4-row blocks."""
    mx, my = 131, 70
    q, aux = _random(mx, my, 7 * -m3 + 10 * order + lim + capa, capa)
    method = (1, order, m3, 0, 0, capa, 3)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3, method=method,
                  mthlim=[lim], mthbc=(1, 1, 2, 2), use_fwave=bool(fw))
    q = O.bc2(p, q)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / my
    ref, cfl_ref = O.dimsp2(p, q, dx, dy, dt, [1.0, 0.9], aux=aux)
    with gpu.Grid2D(SOLVER_VC_ADVECTION, mx, my, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(q)
        cfl = g.dimsp2(dx, dy, dt, method, [lim], (1.0, 0.9), use_fwave=bool(fw))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2], ref[2:-2, 2:-2])


@pytest.mark.parametrize("m3", [-1, -2])
def test_annulus_example_split_matches_reference_run(m3):
    """The annulus example by Godunov / Strang splitting with its capacity
    function, to t = 0.25 through the claw2 loop, against the reference
    build's run (fixture): same steps, same dt, q bit for bit."""
    gold = np.load(os.path.join(GOLD, "annulus_split_40x120.npz"))
    aux, dxc, dyc = O.annulus_aux(40, 120)
    key = f"split{-m3}"
    with gpu.Grid2D(SOLVER_VC_ADVECTION, 40, 120, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(P.annulus_hump(40, 120))
        r = g.claw2(dxc, dyc, 0.0, 0.25, [0.1, 1e9], [1.0, 0.9], 50000, (1, 2, m3, 0, 0, 3, 3),
                    [3], (1, 1, 2, 2))
        qk = g.get_q()
    assert r["info"] == 0 and r["tend"] == float(gold["tend_" + key])
    np.testing.assert_array_equal(r["nv"], gold["nv_" + key])
    np.testing.assert_array_equal(r["dtv"], gold["dtv_" + key])
    np.testing.assert_array_equal(qk[2:-2, 2:-2], gold["q_" + key])


@pytest.mark.parametrize("split", [False, True])
def test_general_path_large_grid_courant_reduction(split):
    """1100 x 700 cells: the records and split-sweep kernels run about 12 k
    waves, whose Courant numbers are stored per wave and reduced by several
    blocks of a second kernel; the step and its Courant number still equal
    the oracle's bit for bit.  The fastest edge velocity sits in one cell
    near the far corner, so a reduction that dropped any block would miss
    it."""
    mx, my = 1100, 700
    q, aux = _random(mx, my, 21, 0)
    aux[my - 3, mx - 7, 0] = 3.0  # the largest |u| of the grid, in the last blocks
    method = (1, 2, -1 if split else 2, 0, 0, 0, 3)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3, method=method,
                  mthlim=[4], mthbc=(1, 1, 2, 2))
    q = O.bc2(p, q)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / my
    if split:
        ref, cfl_ref = O.dimsp2(p, q, dx, dy, dt, [1.0, 0.9], aux=aux)
    else:
        ref, cfl_ref = O.step2(p, q, dx, dy, dt, aux=aux)
    with gpu.Grid2D(SOLVER_VC_ADVECTION, mx, my, [], maux=3) as g:
        g.set_aux(aux)
        g.set_q(q)
        if split:
            cfl = g.dimsp2(dx, dy, dt, method, [4], (1.0, 0.9))
        else:
            cfl = g.step2(dx, dy, dt, method, [4])
        out = g.get_q()
    assert cfl == cfl_ref == dt / dx * 3.0
    np.testing.assert_array_equal(out[2:-2, 2:-2], ref[2:-2, 2:-2])
