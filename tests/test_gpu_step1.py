"""1D path on the GPU against the oracle (oracle/claw1d_oracle.c), bit for
bit: single steps for every solver, limiter, capacity and f-wave case, bc1,
the claw1 loop with retakes, and the two acoustics examples with their
regression frame sums."""
import os

import numpy as np
import pytest

from classic_amd import problems as P
from classic_amd._lib import SOLVER_ACOUSTICS, SOLVER_ADVECTION, SOLVER_VC_ACOUSTICS
from classic_amd.gpu1 import Grid1D
from oracle import oracle1 as O1

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SID = {O1.SOLVER_ACOUSTICS: SOLVER_ACOUSTICS, O1.SOLVER_VC_ACOUSTICS: SOLVER_VC_ACOUSTICS,
       O1.SOLVER_ADVECTION: SOLVER_ADVECTION}


def _params(sid):
    if sid == SOLVER_ACOUSTICS:
        return [1.3, 2.9]
    if sid == SOLVER_ADVECTION:
        return [-0.7]
    return []


@pytest.mark.parametrize("mx", [57, 200, 301])
@pytest.mark.parametrize("case", [(O1.SOLVER_ACOUSTICS, 0, 2, (1, 2), 0),
                                  (O1.SOLVER_ACOUSTICS, 0, 2, (3, 4), 1),
                                  (O1.SOLVER_VC_ACOUSTICS, 2, 2, (4, 2), 0),
                                  (O1.SOLVER_VC_ACOUSTICS, 3, 2, (2, 3), 0),
                                  (O1.SOLVER_ADVECTION, 0, 2, (4,), 0),
                                  (O1.SOLVER_ADVECTION, 0, 1, (0,), 0)])
def test_step1_bitwise(mx, case):
    sid, maux, order, lim, fw = case
    capa = 3 if maux == 3 else 0
    rng = np.random.default_rng(mx + 7 * sid + maux)
    mw = 1 if sid == O1.SOLVER_ADVECTION else 2
    p = O1.Problem1(mx, sid, O1.params1(1.3, 2.9, u=-0.7), meqn=mw, mwaves=mw, maux=maux,
                    method=(1, order, 0, 0, 0, capa, 0), mthlim=lim, use_fwave=bool(fw))
    q = rng.standard_normal((mx + 4, mw))
    aux = None
    if maux:
        aux = np.stack([rng.uniform(0.5, 3, mx + 4), rng.uniform(0.5, 2, mx + 4)] +
                       ([rng.uniform(0.5, 1.5, mx + 4)] if maux > 2 else []), -1)
    dx, dt = 2.0 / mx, 0.3 / mx
    ref, cfl_ref = O1.step1(p, q, dx, dt, aux=aux)
    with Grid1D(SID[sid], mx, _params(SID[sid]), maux=maux) as g:
        if aux is not None:
            g.set_aux(aux)
        g.set_q(q)
        cfl = g.step1(dx, dt, p.method, lim, use_fwave=bool(fw))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2], ref[2:-2])


@pytest.mark.parametrize("mthbc", [(1, 1), (2, 2), (3, 3), (3, 1)])
def test_bc1_bitwise(mthbc):
    rng = np.random.default_rng(2)
    q = rng.standard_normal((74, 2))
    p = O1.Problem1(70, mthbc=mthbc)
    with Grid1D(SOLVER_ACOUSTICS, 70, [1.0, 4.0]) as g:
        g.set_q(q)
        g.bc1(mthbc)
        np.testing.assert_array_equal(g.get_q(), O1.bc1(p, q))


def test_claw1_with_retakes_bitwise():
    rng = np.random.default_rng(4)
    q = rng.standard_normal((104, 2))
    aux = np.stack([rng.uniform(0.5, 3, 104), rng.uniform(0.5, 2, 104)], -1)
    p = O1.Problem1(100, O1.SOLVER_VC_ACOUSTICS, maux=2, mthlim=(2, 3), mthbc=(3, 3))
    r = O1.claw1(p, q, 0.01, 0.0, 0.3, [0.5, 1e9], [1.0, 0.9], 1000, aux=aux)
    with Grid1D(SOLVER_VC_ACOUSTICS, 100, maux=2) as g:
        g.set_aux(aux)
        g.set_q(q)
        rg = g.claw1(0.01, 0.0, 0.3, [0.5, 1e9], [1.0, 0.9], 1000, mthlim=(2, 3), mthbc=(3, 3))
        out = g.get_q()
    assert rg["info"] == r.info == 0 and rg["tend"] == r.tend
    np.testing.assert_array_equal(rg["dtv"], r.dtv)
    np.testing.assert_array_equal(rg["cflv"], r.cflv)
    np.testing.assert_array_equal(rg["nv"], r.nv)
    np.testing.assert_array_equal(out[2:-2], r.q[2:-2])


def test_examples_frames_and_regression_sums():
    g1 = np.load(os.path.join(GOLD, "acoustics1d_example1.npz"))
    q, dx = P.acoustics1d_gaussian(400)
    with Grid1D(SOLVER_ACOUSTICS, 400, [1.0, 4.0]) as g:
        g.set_q(q)
        fr = g.frames(dx, 0.0, 1.0, 2, 0.1, nmax=500, mthlim=(4, 4))
    for k, (t, qq) in enumerate(fr):
        assert t == g1["t"][k]
        np.testing.assert_array_equal(qq[2:-2], g1["q"][k][2:-2])
        np.testing.assert_allclose(qq[2:-2].sum(axis=0), g1["regression_sums"][k], rtol=1e-14,
                                   atol=1e-14)
    g2 = np.load(os.path.join(GOLD, "acoustics1d_heterogeneous.npz"))
    q, aux, dx = P.acoustics1d_heterogeneous(500)
    with Grid1D(SOLVER_VC_ACOUSTICS, 500, maux=2) as g:
        g.set_aux(aux)
        g.set_q(q)
        fr = g.frames(dx, 0.0, 5.0, 2, 1.0, dt_max=1e9, mthlim=(4, 4))
    for k, (t, qq) in enumerate(fr):
        assert t == g2["t"][k]
        np.testing.assert_array_equal(qq[2:-2], g2["q"][k][2:-2])
        np.testing.assert_allclose(qq[2:-2].sum(axis=0), g2["regression_sums"][k], rtol=1e-14,
                                   atol=1e-14)


def test_step1_long_grid_courant_reduction():
    """40,000 cells: 667 strips, whose Courant numbers are stored per strip
    and reduced by several blocks; the fastest sound speed sits in one cell
    near the end, and the step and its Courant number equal the oracle's
    bit for bit."""
    mx = 40000
    rng = np.random.default_rng(9)
    p = O1.Problem1(mx, O1.SOLVER_VC_ACOUSTICS, O1.params1(1.3, 2.9), meqn=2, mwaves=2, maux=2,
                    method=(1, 2, 0, 0, 0, 0, 0), mthlim=(4, 4))
    q = rng.standard_normal((mx + 4, 2))
    aux = np.stack([rng.uniform(0.5, 3, mx + 4), rng.uniform(0.5, 2, mx + 4)], -1)
    aux[mx - 5, 1] = 4.0  # the largest sound speed, in the last strips
    dx, dt = 2.0 / mx, 0.4 / mx
    ref, cfl_ref = O1.step1(p, q, dx, dt, aux=aux)
    with Grid1D(SOLVER_VC_ACOUSTICS, mx, _params(SOLVER_VC_ACOUSTICS), maux=2) as g:
        g.set_aux(aux)
        g.set_q(q)
        cfl = g.step1(dx, dt, p.method, (4, 4))
        out = g.get_q()
    assert cfl == cfl_ref == dt / dx * 4.0
    np.testing.assert_array_equal(out[2:-2], ref[2:-2])
