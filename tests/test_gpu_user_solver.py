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


# ------------------------------------------------------ 3D split step
# user_vc_acoustics3 restates the built-in 3D solver, whose split step equals
# the oracle's dimsp3 bit for bit: the user sweeps must too, Courant numbers,
# dimsp3's early returns and the claw3 dt sequence included.
# user_advection3 on data that varies along one axis only: the two other
# sweeps see no jumps and leave q as it is, and the remaining sweep is the 1D
# oracle's advection step1 on each line (rounded in step1's order, hence a
# tolerance); at Courant number 1 each sweep shifts integer data by one
# cell exactly.

from classic_amd.gpu3 import Grid3D, UserSolver3  # noqa: E402
from oracle import oracle1 as O1  # noqa: E402
from oracle import oracle3 as O3  # noqa: E402
from tests.helpers import max_rel_err3, random_layered  # noqa: E402


@pytest.fixture(scope="module")
def uvc3():
    with UserSolver3(CO, "user_vc_acoustics3") as s:
        yield s


@pytest.fixture(scope="module")
def uadv3():
    with UserSolver3(CO, "user_advection3") as s:
        yield s


def _ugrid3(solver, p, aux, q, params=(), **kw):
    g = Grid3D(p.mx, p.my, p.mz, maux=p.maux, solver=solver, params=params, **kw)
    if aux is not None:
        g.set_aux(aux)
    g.set_q(q)
    return g


def test_user_solver3_shapes(uvc3, uadv3):
    assert (uvc3.meqn, uvc3.mwaves, uvc3.maux, uvc3.arith) == (4, 2, 2, 0)
    assert (uadv3.meqn, uadv3.mwaves, uadv3.maux, uadv3.arith) == (1, 1, 0, 0)
    with UserSolver3(CO_FAST, "user_vc_acoustics3") as s:
        assert s.arith == 1


@pytest.mark.parametrize("order,lim,fw", [(2, (3, 3), False), (2, (4, 1), False),
                                          (2, (2, 3), True), (1, (0, 0), False),
                                          (2, (5, 2), False)])
def test_user_vc_acoustics3_dimsp3_equals_oracle(uvc3, order, lim, fw):
    rng = np.random.default_rng(2 + order + lim[0])
    p = O3.Problem3(70, 9, 6, mthlim=lim, method=(1, order, -1, 0, 0, 0, 0), use_fwave=fw)
    q, aux = random_layered(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt)
    with _ugrid3(uvc3, p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=lim, use_fwave=fw)
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("shape", [(70, 9, 6), (9, 70, 6)])
def test_user_vc_acoustics3_dimsp3_early_return(uvc3, shape):
    """dt above the x (first shape) or y (second) sweep's Courant limit:
    dimsp3 returns after that sweep with its Courant number and state."""
    rng = np.random.default_rng(5)
    p = O3.Problem3(*shape, mthlim=(3, 3), method=(1, 2, -1, 0, 0, 0, 0))
    q, aux = random_layered(p, rng)
    dt = 1.5 * min(p.geom.dx, p.geom.dy) / 2.0
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt, cflv=(1.0, 0.9))
    with _ugrid3(uvc3, p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=(3, 3), cflv=(1.0, 0.9))
        out = g.get_q()
    assert cfl > 1.0 and cfl == cfl_ref
    np.testing.assert_array_equal(out, ref)


def test_user_vc_acoustics3_claw3_with_retakes_equals_oracle(uvc3):
    from classic_amd import problems as P
    p = O3.Problem3(36, 28, 20, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2),
                    method=(1, 2, -1, 0, 0, 0, 0))
    qc, _, _ = P.acoustics3d_pulse(36, seed=4, width=0.3)
    q = qc[: p.shape()[0], : p.shape()[1]].copy()
    aux = P.layered_medium_aux(36, 28, 20, below=(1.0, 1.0), above=(6.0, 2.0))
    r = O3.claw3(p, q, aux, 0.0, 0.4, [0.2, 1e9], [1.0, 0.9], 500)
    with _ugrid3(uvc3, p, aux, q) as g:
        rg = g.claw3(0.0, 0.4, [0.2, 1e9], [1.0, 0.9], 500, method=p.method, mthlim=(2, 2),
                     mthbc=(3, 3, 3, 3, 3, 3))
        out = g.get_q()
    assert rg.info == r.info == 0
    assert rg.tend == r.tend and r.nv[1] > 2
    np.testing.assert_array_equal(rg.nv, r.nv)
    np.testing.assert_array_equal(rg.dtv, r.dtv)
    np.testing.assert_array_equal(rg.cflv, r.cflv)
    np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], r.q[2:-2, 2:-2, 2:-2])


def test_user_vc_acoustics3_device_b4step3_equals_builtin(uvc3):
    """claw3 with a device b4step3 that raises every sound speed each attempt
    (retakes) and walls in y: the user solver reads the changed aux itself and
    gives the built-in split step's results, dt sequence and Courant numbers
    bit for bit."""
    from tests import hooks_cases as H
    from tests.test_gpu_hooks import _dev_hooks
    q0, aux = H.initial3("acoustics3d_src_strang_split")
    c = H.CASES3["acoustics3d_src_strang_split"]
    res = []
    for solver in (None, uvc3):
        with _dev_hooks(0.25, 400, b4step2="test_b4step3_speed") as dh, \
                Grid3D(c["mx"], c["my"], c["mz"], maux=2, solver=solver) as g:
            g.set_aux(aux)
            g.set_q(q0)
            out = g.claw3(0.0, 0.3, [0.3, 1e9], [1.0, 0.9], 500, method=(1, 2, -1, 0, 0, 0, 0),
                          mthlim=(3, 3), mthbc=(1, 1, 3, 3, 1, 1), dev_hooks=dh)
            res.append((g.get_q()[2:-2, 2:-2, 2:-2], out))
    (qb, ob), (qu, ou) = res
    assert ob.nv[1] == ou.nv[1] and ob.tend == ou.tend
    np.testing.assert_array_equal(ou.dtv[2:5], ob.dtv[2:5])
    np.testing.assert_array_equal(ou.cflv[2:4], ob.cflv[2:4])
    np.testing.assert_array_equal(qu, qb)


def test_user_vc_acoustics3_fast_within_tolerance():
    rng = np.random.default_rng(8)
    p = O3.Problem3(70, 9, 6, mthlim=(4, 4), method=(1, 2, -1, 0, 0, 0, 0))
    q, aux = random_layered(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt)
    with UserSolver3(CO_FAST, "user_vc_acoustics3") as s, _ugrid3(s, p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=(4, 4))
        out = g.get_q()
    assert abs(cfl - cfl_ref) <= 1e-15 * cfl_ref
    assert max_rel_err3(out, ref) <= 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("order,lim", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (1, 0)])
def test_user_advection3_sweep_equals_1d_oracle(uadv3, axis, order, lim):
    n3 = [9, 7, 6]
    n3[axis] = 70
    vel = [-0.8, 0.6, 0.9]
    vel[axis] = 0.7 if lim % 2 else -0.7
    p = O3.Problem3(*n3, maux=0, mthlim=(lim,), method=(1, order, -1, 0, 0, 0, 0))
    h = (2.0 / n3[0], 2.0 / n3[1], 2.0 / n3[2])
    rng = np.random.default_rng(10 * axis + lim)
    line = rng.standard_normal(n3[axis] + 4)
    shape = [n + 4 for n in n3]  # x, y, z extents
    q = np.broadcast_to(line.reshape([-1 if a == axis else 1 for a in (2, 1, 0)]),
                        (shape[2], shape[1], shape[0])).copy()[..., None]
    dt = 0.8 * h[axis] / 0.7
    with _ugrid3(uadv3, p, None, q, params=vel) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=(lim,))
        out = g.get_q()[..., 0]
    p1 = O1.Problem1(n3[axis], O1.SOLVER_ADVECTION, O1.params1(u=vel[axis]), meqn=1, mwaves=1,
                     method=(1, order, 0, 0, 0, 0, 0), mthlim=(lim,))
    ref, cfl1 = O1.step1(p1, line.reshape(-1, 1), h[axis], dt)
    assert cfl == max(dt / h[a] * abs(vel[a]) for a in range(3))
    assert cfl1 == dt / h[axis] * abs(vel[axis])
    # the cells 1..n of the interior lines (step1 also adds fluctuations to
    # the ghost cells 0 and n+1, which step3ds leaves alone)
    lines = np.moveaxis(out, 2 - axis, -1)[2:-2, 2:-2, 2:-2]
    assert np.max(np.abs(lines - ref[2:-2, 0])) <= 1e-13


def test_user_advection3_courant_one_shifts_exactly(uadv3):
    """u, v, w = +h/dt, -h/dt, +h/dt and periodic boundaries: four claw3
    steps shift integer data by (4, -4, 4) cells exactly."""
    n3 = (70, 9, 6)
    h = 1.0 / 16.0
    lower = (0.0, 0.0, 0.0)
    upper = tuple(n * h for n in n3)
    p = O3.Problem3(*n3, maux=0, mthlim=(3,), method=(0, 2, -1, 0, 0, 0, 0))
    rng = np.random.default_rng(6)
    q = rng.integers(-8, 9, (n3[2] + 4, n3[1] + 4, n3[0] + 4, 1)).astype(np.float64)
    with _ugrid3(uadv3, p, None, q, params=(1.0, -1.0, 1.0), lower=lower, upper=upper) as g:
        r = g.claw3(0.0, 4 * h, [h, 1e9], [1.0, 0.9], 10, method=p.method, mthlim=(3,),
                    mthbc=(2, 2, 2, 2, 2, 2))
        out = g.get_q()[2:-2, 2:-2, 2:-2, 0]
    assert r.info == 0 and r.nv[1] == 4 and r.cflv[2] == 1.0
    want = np.roll(q[2:-2, 2:-2, 2:-2, 0], (4, -4, 4), axis=(0, 1, 2))
    np.testing.assert_array_equal(out, want)


def test_user_solver3_refusals(uvc3, uadv3):
    p = O3.Problem3(12, 10, 8, mthlim=(3, 3), method=(1, 2, -1, 0, 0, 0, 0))
    q, aux = random_layered(p, np.random.default_rng(1))
    with _ugrid3(uvc3, p, aux, q) as g:
        with pytest.raises(ClawGpuError) as e:
            g.step3(0.01)                                # unsplit step: not for user solvers
        assert e.value.code == ERR_UNSUPPORTED
        with pytest.raises(ClawGpuError) as e:
            g.dimsp3(0.01, method=(1, 2, -1, 0, 0, 1, 2))  # capacity function
        assert e.value.code == ERR_UNSUPPORTED
        with pytest.raises(ClawGpuError) as e:
            g.bc3((1, 1, 1, 1, 1, 0))                    # piston
        assert e.value.code == ERR_UNSUPPORTED
        with pytest.raises(ClawGpuError) as e:
            g.set_arith(1)                               # built EXACT
        assert e.value.code == ERR_UNSUPPORTED
        with pytest.raises(ClawGpuError) as e:
            g.time_steps(0.01, 2, method=(1, 2, -1, 0, 0, 0, 2))
        assert e.value.code == ERR_UNSUPPORTED
        g.set_arith(0)
    pa = O3.Problem3(12, 10, 8, maux=0, mthlim=(3,), method=(1, 2, -1, 0, 0, 0, 0))
    with _ugrid3(uadv3, pa, None, np.zeros(pa.shape(1)), params=(1.0, 1.0, 1.0)) as g:
        g.bc3((2, 2, 1, 1, 2, 2))                        # periodic and extrapolation: fine
        with pytest.raises(ClawGpuError) as e:
            g.bc3((3, 1, 1, 1, 1, 1))                    # a wall negates component 2: none
        assert e.value.code == ERR_ARG
    with pytest.raises(ClawGpuError) as e:
        Grid3D(8, 8, 8, maux=1, solver=uvc3)             # aux planes the solver reads
    assert e.value.code == ERR_ARG
    with pytest.raises(ClawGpuError) as e:
        UserSolver3(CO, "user_acoustics")                # a 2D solver
    assert e.value.code == ERR_ARG
    with pytest.raises(ClawGpuError) as e:
        Grid3D(8, 8, 8, maux=2, solver=uvc3, params=np.zeros(17))
    assert e.value.code == ERR_ARG


def test_user_vc_acoustics3_courant_reduction_planted_maximum(uvc3):
    """130 x 40 x 30 cells (about 4,500 x-sweep waves, their Courant numbers
    reduced by several blocks) with the fastest sound speed in one cell
    near the far corner: dimsp3's Courant number and state equal the
    oracle's bit for bit."""
    rng = np.random.default_rng(12)
    p = O3.Problem3(130, 40, 30, mthlim=(3, 3), method=(1, 2, -1, 0, 0, 0, 0))
    q, aux = random_layered(p, rng)
    aux[30, 39, 125, 1] = 2.5  # (k, j, i) = (29, 38, 124): the largest c of the grid
    dt = 0.3 * p.geom.dx / 2.5
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt)
    with _ugrid3(uvc3, p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=(3, 3))
        out = g.get_q()
    assert cfl == cfl_ref == dt / p.geom.dx * 2.5
    np.testing.assert_array_equal(out, ref)
