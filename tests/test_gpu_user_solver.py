"""User Riemann solvers (include/clawgpu_solver.h): the policies of
tests/user_solvers.hip, built into gfx950 code objects by
__graft_entry__.build() and loaded with clawgpu_solver2_load, run every 2D
path of the library (unsplit step, split steps, the device-side claw2 loop,
the host loop with user routines, the solver probe) through the library's
kernels instantiated with the user's policy.

* user_acoustics restates the built-in acoustics solver, so in EXACT
  arithmetic each path must give the built-in solver's results bit for bit,
  and in FAST arithmetic agree to 1e-12.
* user_advection is constant-velocity advection, which is the reference's
  variable-coefficient advection solver (rpn2/rpt2_vc_advection of the
  annulus example) with constant edge velocities: each path must give the
  oracle's vc-advection results bit for bit, Courant numbers and the claw2
  step/retake sequence included."""
import ctypes as C
import os

import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, gpu
from classic_amd._lib import ERR_ARG, ERR_UNSUPPORTED, ClawGpuError
from oracle import oracle as O
from tests.helpers import max_rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CO = os.path.join(HERE, "user_solvers.co")
CO_FAST = os.path.join(HERE, "user_solvers_fast.co")
RHO, BULK = 1.3, 2.9


def _zc():
    """(Z, c) as the library derives them from (rho, bulk)."""
    c = np.sqrt(BULK / RHO)
    return [RHO * c, c]


def _acoustic_state(mx, my, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(-1.0, 1.0, mx + 4)[None, :]
    y = np.linspace(-1.0, 1.0, my + 4)[:, None]
    q = np.zeros((my + 4, mx + 4, 3))
    q[..., 0] = np.exp(-12.0 * ((x - 0.2) ** 2 + (y + 0.1) ** 2))
    q += 0.05 * rng.standard_normal(q.shape)
    return q


def _both(solver, fn, arith=0, params=None):
    """fn(grid) for the built-in acoustics solver, then for `solver`."""
    out = []
    for sid, p in ((SOLVER_ACOUSTICS, [RHO, BULK]), (solver, _zc() if params is None else params)):
        g = gpu.Grid2D(sid, fn.mx, fn.my, p)
        try:
            if arith and sid == SOLVER_ACOUSTICS:
                g.set_arith(arith)
            out.append(fn(g))
        finally:
            g.close()
    return out


def _sized(mx, my):
    def wrap(f):
        f.mx, f.my = mx, my
        return f
    return wrap


@pytest.fixture(scope="module")
def ua():
    with gpu.UserSolver(CO, "user_acoustics") as s:
        yield s


@pytest.fixture(scope="module")
def uadv():
    with gpu.UserSolver(CO, "user_advection") as s:
        yield s


def test_loaded_shapes(ua, uadv):
    assert (ua.meqn, ua.mwaves, ua.arith) == (3, 2, 0)
    assert (uadv.meqn, uadv.mwaves, uadv.arith) == (1, 1, 0)
    with gpu.UserSolver(CO_FAST, "user_acoustics") as f:
        assert (f.meqn, f.mwaves, f.arith) == (3, 2, 1)


@pytest.mark.parametrize("order,trans,lim,fw", [(2, 2, 4, 0), (2, 1, 3, 0), (2, 0, 1, 0),
                                                (1, 2, 0, 0), (2, 2, 2, 1), (2, 2, 5, 0)])
def test_user_acoustics_step2_equals_builtin(ua, order, trans, lim, fw):
    q = _acoustic_state(133, 71, 10 * order + trans + lim)
    dx, dy = 2.0 / 133, 2.0 / 71
    dt = 0.8 * min(dx, dy) / _zc()[1]

    @_sized(133, 71)
    def run(g):
        g.set_q(q)
        g.bc2((1, 3, 2, 2))
        cfl = g.step2(dx, dy, dt, (1, order, trans, 0, 0, 0, 0), [lim, lim], use_fwave=bool(fw))
        return cfl, g.get_q()

    (c0, q0), (c1, q1) = _both(ua, run)
    assert c0 == c1
    np.testing.assert_array_equal(q0, q1)


@pytest.mark.parametrize("m3", [-1, -2])
@pytest.mark.parametrize("order,lim", [(2, 4), (2, 1), (1, 0)])
def test_user_acoustics_dimsp2_equals_builtin(ua, m3, order, lim):
    q = _acoustic_state(97, 130, 3 * order + lim - m3)
    dx, dy = 2.0 / 97, 2.0 / 130
    dt = 0.7 * min(dx, dy) / _zc()[1]

    @_sized(97, 130)
    def run(g):
        g.set_q(q)
        g.bc2((3, 3, 1, 1))
        cfl = g.dimsp2(dx, dy, dt, (1, order, m3, 0, 0, 0, 0), [lim, lim])
        return cfl, g.get_q()

    (c0, q0), (c1, q1) = _both(ua, run)
    assert c0 == c1
    np.testing.assert_array_equal(q0, q1)


@pytest.mark.parametrize("mthbc", [(3, 3, 3, 3), (2, 2, 1, 1)])
def test_user_acoustics_claw2_equals_builtin(ua, mthbc):
    """The device-side claw2 loop (with and without walls) and a retake:
    the first dt gives a Courant number above cflv(1)."""
    q = _acoustic_state(150, 90, 5)
    dx, dy = 2.0 / 150, 2.0 / 90

    @_sized(150, 90)
    def run(g):
        g.set_q(q)
        r = g.claw2(dx, dy, 0.0, 0.35, [0.05, 1e9], [1.0, 0.9], 1000, (1, 2, 2, 0, 0, 0, 0),
                    [4, 4], mthbc)
        return r, g.get_q()

    (r0, q0), (r1, q1) = _both(ua, run)
    assert r0["info"] == r1["info"] == 0 and r0["tend"] == r1["tend"]
    np.testing.assert_array_equal(r0["nv"], r1["nv"])
    np.testing.assert_array_equal(r0["dtv"], r1["dtv"])
    np.testing.assert_array_equal(r0["cflv"], r1["cflv"])
    np.testing.assert_array_equal(q0[2:-2, 2:-2], q1[2:-2, 2:-2])


def test_user_acoustics_probe_equals_builtin(ua):
    rng = np.random.default_rng(4)
    ql, qr, asdq = (rng.standard_normal((300, 3)) for _ in range(3))
    for ixy in (1, 2):
        @_sized(8, 8)
        def run(g):
            return g.solver_probe(ixy, ql, qr, asdq)

        a, b = _both(ua, run)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_user_acoustics_fast_within_1e12_of_builtin_fast():
    q = _acoustic_state(200, 120, 8)
    dx, dy = 2.0 / 200, 2.0 / 120

    @_sized(200, 120)
    def run(g):
        g.set_q(q)
        r = g.claw2(dx, dy, 0.0, 0.3, [0.004, 1e9], [1.0, 0.9], 1000, (1, 2, 2, 0, 0, 0, 0),
                    [4, 4], (3, 3, 1, 1))
        return r, g.get_q()

    with gpu.UserSolver(CO_FAST, "user_acoustics") as f:
        (r0, q0), (r1, q1) = _both(f, run, arith=1)
    assert r0["nv"][1] == r1["nv"][1] and r0["tend"] == r1["tend"]
    err = max_rel_err(q1, q0)
    print("FAST user vs built-in FAST, max rel err", err, flush=True)
    assert err <= 1e-12


def _advection(mx, my, u, v, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((my + 4, mx + 4, 1))
    aux = np.empty((my + 4, mx + 4, 2))
    aux[..., 0] = u
    aux[..., 1] = v
    return q, aux


@pytest.mark.parametrize("uv", [(0.7, -0.4), (-1.1, 0.6)])
@pytest.mark.parametrize("order,trans,lim,fw", [(2, 2, 4, 0), (2, 1, 3, 0), (2, 0, 1, 0),
                                                (1, 2, 0, 0), (1, 1, 0, 0), (2, 2, 2, 1)])
def test_user_advection_step2_equals_oracle_vc_advection(uadv, uv, order, trans, lim, fw):
    mx, my = 131, 70
    q, aux = _advection(mx, my, *uv, seed=order + 3 * trans + 7 * lim)
    method = (1, order, trans, 0, 0, 0, 0)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=2, method=method,
                  mthlim=[lim], mthbc=(1, 1, 2, 2), use_fwave=bool(fw))
    q = O.bc2(p, q)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.3 / my
    ref, cfl_ref = O.step2(p, q, dx, dy, dt, aux=aux)
    with gpu.Grid2D(uadv, mx, my, uv) as g:
        g.set_q(q)
        cfl = g.step2(dx, dy, dt, method, [lim], use_fwave=bool(fw))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2], ref[2:-2, 2:-2])


@pytest.mark.parametrize("m3", [-1, -2])
@pytest.mark.parametrize("order,lim", [(2, 4), (2, 3), (1, 0)])
def test_user_advection_dimsp2_equals_oracle(uadv, m3, order, lim):
    mx, my = 131, 70
    q, aux = _advection(mx, my, -0.8, 0.5, seed=11 * order + lim - m3)
    method = (1, order, m3, 0, 0, 0, 0)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=2, method=method,
                  mthlim=[lim], mthbc=(1, 1, 2, 2))
    q = O.bc2(p, q)
    dx, dy, dt = 1.0 / mx, 1.0 / my, 0.3 / my
    ref, cfl_ref = O.dimsp2(p, q, dx, dy, dt, [1.0, 0.9], aux=aux)
    with gpu.Grid2D(uadv, mx, my, (-0.8, 0.5)) as g:
        g.set_q(q)
        cfl = g.dimsp2(dx, dy, dt, method, [lim], (1.0, 0.9))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2], ref[2:-2, 2:-2])


def test_user_advection_claw2_with_retakes_equals_oracle(uadv):
    mx, my = 64, 48
    q, aux = _advection(mx, my, 0.9, -0.6, seed=3)
    method = (1, 2, 2, 0, 0, 0, 0)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=2, method=method,
                  mthlim=[4], mthbc=(2, 2, 1, 1))
    r = O.claw2(p, q, 1.0 / mx, 1.0 / my, 0.0, 0.4, [0.1, 1e9], [1.0, 0.9], 1000, aux=aux)
    with gpu.Grid2D(uadv, mx, my, (0.9, -0.6)) as g:
        g.set_q(q)
        rg = g.claw2(1.0 / mx, 1.0 / my, 0.0, 0.4, [0.1, 1e9], [1.0, 0.9], 1000, method, [4],
                     (2, 2, 1, 1))
        out = g.get_q()
    assert len(r.dt_trace) > r.nv[1]  # the run had retakes
    assert rg["info"] == r.info == 0 and rg["tend"] == r.tend
    np.testing.assert_array_equal(rg["dtv"], r.dtv)
    np.testing.assert_array_equal(rg["cflv"], r.cflv)
    np.testing.assert_array_equal(rg["nv"], r.nv)
    np.testing.assert_array_equal(out[2:-2, 2:-2], r.q[2:-2, 2:-2])


def test_user_solver_with_host_routines(uadv):
    """The host loop with a user src2 (decay) on the user solver gives the
    oracle's claw2 followed by the same source calls: here checked by the
    source's effect alone, a uniform field decaying as (1 - k dt) per step."""
    mx, my = 40, 30
    q = np.ones((my + 4, mx + 4, 1))
    dts = []

    def src2(qv, aux, t, dt):
        dts.append(dt)
        qv[...] -= (0.5 * dt) * qv

    with gpu.Grid2D(uadv, mx, my, (0.3, 0.2)) as g:
        g.set_q(q)
        r = g.claw2(1.0 / mx, 1.0 / my, 0.0, 0.2, [0.01, 1e9], [1.0, 0.9], 100,
                    (0, 2, 2, 0, 1, 0, 0), [4], (2, 2, 2, 2), src2=src2)
        out = g.get_q()
    assert r["tend"] == pytest.approx(0.2) and len(dts) == r["nv"][1] >= 20
    want = 1.0
    for dt in dts:
        want -= (0.5 * dt) * want
    np.testing.assert_array_equal(out[2:-2, 2:-2, 0], np.full((my, mx), want))


def test_user_solver_errors(uadv, ua):
    with pytest.raises(ClawGpuError, match="not found"):
        gpu.UserSolver(CO, "no_such_solver")
    with pytest.raises(ClawGpuError):
        gpu.UserSolver(CO + ".missing", "user_acoustics")
    with pytest.raises(ClawGpuError) as e:
        gpu.Grid2D(uadv, 10, 10, np.zeros(17))
    assert e.value.code == ERR_ARG
    with gpu.Grid2D(uadv, 20, 20, (0.5, 0.5)) as g:
        g.set_q(np.zeros((24, 24, 1)))
        with pytest.raises(ClawGpuError, match="meqn >= 3") as e:
            g.bc2((3, 3, 1, 1))
        assert e.value.code == ERR_ARG
        with pytest.raises(ClawGpuError) as e:
            g.set_arith(1)
        assert e.value.code == ERR_UNSUPPORTED
        params = np.zeros(1281, dtype=np.float32)
        rc = g.lib.clawgpu_learned_set_params(g.h, params.ctypes.data_as(C.POINTER(C.c_float)),
                                              params.size)
        assert rc == ERR_UNSUPPORTED
        # the errors above leave the context usable
        g.bc2((2, 2, 2, 2))
        assert g.step2(0.05, 0.05, 0.01, (1, 2, 2, 0, 0, 0, 0), [4]) == pytest.approx(0.1)
