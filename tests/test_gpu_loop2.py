"""The device-side claw2 loop (clawgpu_claw2 without user routines): every
attempt of a step is one kernel launch that decides the previous attempt on
the device (src/2d/claw2.f:695-753) and reads the ghost cells through the
bc2 index maps (src/2d/bc2.f:4-215) instead of a ghost-fill pass.  Checked
against the oracle's claw2 (bc2 + copyq2 + step2 + retake logic), for every
boundary-condition code on each side, with rejected steps, runs that end on
tend and on the step count, and the non-finite-speed stop."""
import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from classic_amd import problems as P
from classic_amd._lib import ClawGpuError
from oracle import oracle as O
from tests.helpers import max_rel_err, smooth_acoustics

pytestmark = pytest.mark.gpu

BCS = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (3, 1, 2, 2), (1, 3, 3, 1), (2, 2, 3, 1),
       (3, 3, 2, 2)]


def _case(solver, mx, my, mthbc, method=(1, 2, 2, 0, 0, 0, 0)):
    if solver == "acoustics":
        q = smooth_acoustics(mx, my, seed=7)
        dx, dy = 1.0 / mx, 1.0 / my
        p = O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                      method=method, mthlim=[4, 2], mthbc=mthbc)
        return q, dx, dy, p, (SOLVER_ACOUSTICS, [1.0, 4.0]), [4, 2]
    q, dx, dy = P.euler_quadrants(mx, my, seed=3)
    p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                  method=method, mthlim=[4, 4, 4], mthbc=mthbc)
    return q, dx, dy, p, (SOLVER_EULER, [1.4, 1.0]), [4, 4, 4]


@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("mthbc", BCS)
@pytest.mark.parametrize("solver", ["acoustics", "euler"])
def test_device_loop_all_bcs(solver, mthbc, arith):
    mx, my = 131, 53  # three strips, the last one ragged
    q, dx, dy, p, (sid, par), mthlim = _case(solver, mx, my, mthbc)
    # a first dt five times too large: rejected, then retaken (claw2.f:699-714)
    dtv = [5.0 * dx, 1e99]
    r = O.claw2(p, q, dx, dy, 0.0, 1e9, dtv, [1.0, 0.9], 14)
    with gpu.Grid2D(sid, mx, my, par, arith=arith) as g:
        g.set_q(q)
        out = g.claw2(dx, dy, 0.0, 1e9, dtv, [1.0, 0.9], 14, p.method, mthlim, mthbc)
        qg = g.get_q()
    assert out["info"] == r.info
    assert out["nv"][1] == r.nv[1] == 14
    if arith == 0:
        assert np.array_equal(qg[2:-2, 2:-2], r.q[2:-2, 2:-2])
        np.testing.assert_array_equal(out["dtv"], r.dtv)
        np.testing.assert_array_equal(out["cflv"], r.cflv)
        assert out["tend"] == r.tend
    else:
        assert max_rel_err(qg, r.q) <= 1e-12
        np.testing.assert_allclose(out["dtv"], r.dtv, rtol=1e-12)
        np.testing.assert_allclose(out["cflv"], r.cflv, rtol=1e-12)


def test_device_loop_ends_on_tend_in_chunks():
    """Many steps (several read-back chunks), dt clamped to hit tend exactly
    (claw2.f:601); the run stops on t >= tend with steps to spare."""
    q, dx, dy, p, (sid, par), mthlim = _case("acoustics", 64, 48, (3, 3, 3, 3))
    tend = 0.8
    r = O.claw2(p, q, dx, dy, 0.0, tend, [1e-3, 1e99], [1.0, 0.9], 5000)
    with gpu.Grid2D(sid, 64, 48, par) as g:
        g.set_q(q)
        out = g.claw2(dx, dy, 0.0, tend, [1e-3, 1e99], [1.0, 0.9], 5000, p.method, mthlim,
                      (3, 3, 3, 3))
        qg = g.get_q()
    assert r.nv[1] > 64
    assert out["info"] == r.info == 0
    assert out["nv"][1] == r.nv[1]
    assert out["tend"] == r.tend == tend
    np.testing.assert_array_equal(out["dtv"], r.dtv)
    assert np.array_equal(qg[2:-2, 2:-2], r.q[2:-2, 2:-2])


def test_device_loop_back_to_back_calls():
    """Two claw2 calls in a row continue from the first call's buffer and dt,
    as the frame loop of claw2ez does."""
    q, dx, dy, p, (sid, par), mthlim = _case("euler", 90, 70, (1, 1, 1, 1))
    r1 = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 9)
    r2 = O.claw2(p, r1.q, dx, dy, r1.tend, 1e9, [r1.dtv[4], 1e99], [1.0, 0.9], 7)
    with gpu.Grid2D(sid, 90, 70, par) as g:
        g.set_q(q)
        o1 = g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 9, p.method, mthlim, (1, 1, 1, 1))
        o2 = g.claw2(dx, dy, o1["tend"], 1e9, [o1["dtv"][4], 1e99], [1.0, 0.9], 7, p.method,
                     mthlim, (1, 1, 1, 1))
        qg = g.get_q()
    assert o2["nv"][1] == r2.nv[1] == 7
    np.testing.assert_array_equal(o2["dtv"], r2.dtv)
    assert np.array_equal(qg[2:-2, 2:-2], r2.q[2:-2, 2:-2])


def test_device_loop_fixed_dt_info12_and_zero_steps():
    q, dx, dy, p, (sid, par), mthlim = _case("euler", 40, 40, (1, 1, 1, 1),
                                             method=(0, 2, 2, 0, 0, 0, 0))
    r = O.claw2(p, q, dx, dy, 0.0, 50 * dx, [dx, 1e99], [1.0, 0.9], 500)
    with gpu.Grid2D(sid, 40, 40, par) as g:
        g.set_q(q)
        out = g.claw2(dx, dy, 0.0, 50 * dx, [dx, 1e99], [1.0, 0.9], 500, p.method, mthlim,
                      (1, 1, 1, 1))
        out0 = g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 0, (1, 2, 2, 0, 0, 0, 0),
                       mthlim, (1, 1, 1, 1))
    assert out["info"] == r.info == 12
    np.testing.assert_array_equal(out["cflv"], r.cflv)
    # no steps asked for: t stays below tend with nv = maxn, info 11 as in claw2.f:766-770
    p0 = O.Problem(4, 3, 40, 40, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                   mthlim=mthlim, mthbc=(1, 1, 1, 1))
    r0 = O.claw2(p0, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 0)
    assert out0["nv"][1] == r0.nv[1] == 0
    assert out0["info"] == r0.info == 11


def test_device_loop_nonfinite_speed_stops():
    """A state the Roe solver cannot handle (negative pressure after a few
    steps with a reflecting wall, as in the oracle's own test) stops the
    device loop with info = -5 and an error return, where the reference
    would retake forever."""
    q, dx, dy = P.euler_quadrants(45, 38, seed=0)
    r = O.claw2(O.Problem(4, 3, 45, 38, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                          mthlim=[4, 4, 4], mthbc=(1, 3, 2, 2)),
                q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 15)
    assert r.info == -5
    with gpu.Grid2D(SOLVER_EULER, 45, 38, [1.4, 1.0]) as g:
        g.set_q(q)
        with pytest.raises(ClawGpuError):
            g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 15, (1, 2, 2, 0, 0, 0, 0),
                    [4, 4, 4], (1, 3, 2, 2))
