"""User routines of the plug-in contract (bc, b4step, src) on the device
claw1 / claw2 / claw3 loops (clawgpu1_claw1_hooks, clawgpu_claw2_hooks,
clawgpu3_claw3_hooks) against the reference loops run with the same
routines (tests/golden/claw{1,2,3}_hooks_*.npz, written by the reference
build through tests/golden/make_golden_hooks.py): the same calls in the same
order with the same (t, dt) — retakes and Strang half steps included — and
the same final state, bit for bit in 1D and 2D (EXACT arithmetic) and for the
split 3D step, to 1e-12 for the unsplit 3D step."""
import os

import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, SOLVER_VC_ACOUSTICS, gpu
from classic_amd.gpu1 import Grid1D
from classic_amd.gpu3 import Grid3D
from tests import hooks_cases as H
from tests.helpers import interior, max_rel_err3

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", list(H.CASES))
def test_hooks_match_reference_claw2(name):
    c = H.CASES[name]
    g0 = np.load(os.path.join(GOLD, f"claw2_hooks_{name}.npz"))
    q0, dx, dy = H.initial(name)
    np.testing.assert_array_equal(q0, g0["q0"])
    euler = c["solver"] == "euler"
    sid, params, mthlim = ((SOLVER_EULER, [1.4, 1.0], [4, 4, 4]) if euler
                           else (SOLVER_ACOUSTICS, [1.0, 4.0], [4, 4]))
    b4, src, log = H.python_hooks(c["k"])
    with gpu.Grid2D(sid, c["mx"], c["my"], params) as g:
        g.set_q(q0)
        out = g.claw2(dx, dy, 0.0, c["tend"], list(c["dtv"]), [1.0, 0.9], 500, c["method"], mthlim,
                      c["mthbc"], b4step2=b4, src2=src)
        q = g.get_q()
    print(name, "steps", out["nv"][1], "calls", len(log), flush=True)
    np.testing.assert_array_equal(np.array(log), g0["log"])
    assert out["info"] == int(g0["info"]) and out["tend"] == float(g0["tend"])
    assert out["nv"][1] == g0["nv"][1]
    np.testing.assert_array_equal(out["dtv"][2:5], g0["dtv"][2:5])
    np.testing.assert_array_equal(out["cflv"][2:4], g0["cflv"][2:4])
    np.testing.assert_array_equal(interior(q), interior(g0["q"]))


def _extrap(q, aux, t, dt, mthbc, mbc=2):
    """bc2's extrapolation (bc2.f:4-215, mthbc = 1) written as a user
    routine: ghost columns first, then ghost rows over every column."""
    q[:, :mbc] = q[:, mbc:mbc + 1]
    q[:, -mbc:] = q[:, -mbc - 1:-mbc]
    q[:mbc, :] = q[mbc:mbc + 1, :]
    q[-mbc:, :] = q[-mbc - 1:-mbc, :]


def test_user_bc2_equals_builtin():
    """A user boundary routine (called on the host copy of q before every
    step attempt) that extrapolates gives the built-in mthbc = 1 result bit
    for bit, retakes included."""
    q0, dx, dy = H.initial("euler_src_strang")
    method = (1, 2, 2, 0, 0, 0, 0)
    res = []
    for user in (False, True):
        with gpu.Grid2D(SOLVER_EULER, q0.shape[1] - 4, q0.shape[0] - 4, [1.4, 1.0]) as g:
            g.set_q(q0)
            kw = dict(bc2=_extrap) if user else {}
            out = g.claw2(dx, dy, 0.0, 0.1, [0.02, 1e99], [1.0, 0.9], 500, method, [4, 4, 4],
                          (1, 1, 1, 1) if not user else (0, 0, 0, 0), **kw)
            res.append((interior(g.get_q()), out))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1]["nv"][1] == res[1][1]["nv"][1] and res[0][1]["tend"] == res[1][1]["tend"]


def test_hook_errors_surface():
    q0, dx, dy = H.initial("acoustics_src_godunov")

    def bad(q, aux, t, dt):
        raise ValueError("boom")

    with gpu.Grid2D(SOLVER_ACOUSTICS, 30, 24, [1.0, 4.0]) as g:
        g.set_q(q0)
        with pytest.raises(ValueError, match="boom"):
            g.claw2(dx, dy, 0.0, 0.01, [1e-3, 1e99], [1.0, 0.9], 5, (1, 2, 2, 0, 1, 0, 0), [4, 4],
                    (3, 3, 3, 3), src2=bad)


@pytest.mark.parametrize("name", list(H.CASES1))
def test_hooks_match_reference_claw1(name):
    c = H.CASES1[name]
    g0 = np.load(os.path.join(GOLD, f"claw1_hooks_{name}.npz"))
    q0, aux, dx = H.initial1(name)
    np.testing.assert_array_equal(q0, g0["q0"])
    vc = c["solver"] == "vc_acoustics"
    b4, src, log = H.python_hooks(c["k"], dim=1)
    with Grid1D(SOLVER_VC_ACOUSTICS if vc else SOLVER_ACOUSTICS, c["mx"], [] if vc else [1.3, 2.9],
                maux=c["maux"]) as g:
        if aux is not None:
            g.set_aux(aux)
        g.set_q(q0)
        out = g.claw1(dx, 0.0, c["tend"], list(c["dtv"]), [1.0, 0.9], 500, c["method"],
                      c["mthlim"], c["mthbc"], b4step1=b4, src1=src)
        q = g.get_q()
    print(name, "steps", out["nv"][1], "calls", len(log), flush=True)
    np.testing.assert_array_equal(np.array(log), g0["log"])
    assert out["info"] == int(g0["info"]) and out["tend"] == float(g0["tend"])
    assert out["nv"][1] == g0["nv"][1]
    np.testing.assert_array_equal(out["dtv"][2:5], g0["dtv"][2:5])
    np.testing.assert_array_equal(out["cflv"][2:4], g0["cflv"][2:4])
    np.testing.assert_array_equal(q[2:-2], g0["q"][2:-2])


@pytest.mark.parametrize("name", list(H.CASES3))
def test_hooks_match_reference_claw3(name):
    c = H.CASES3[name]
    g0 = np.load(os.path.join(GOLD, f"claw3_hooks_{name}.npz"))
    q0, aux = H.initial3(name)
    np.testing.assert_array_equal(q0, g0["q0"])
    b4, src, log = H.python_hooks(c["k"], dim=3)
    with Grid3D(c["mx"], c["my"], c["mz"], maux=2) as g:
        g.set_aux(aux)
        g.set_q(q0)
        out = g.claw3(0.0, c["tend"], list(c["dtv"]), [1.0, 0.9], 500, method=c["method"],
                      mthlim=c["mthlim"], mthbc=c["mthbc"], b4step3=b4, src3=src)
        q = g.get_q()
    print(name, "steps", out.nv[1], "calls", len(log), flush=True)
    np.testing.assert_array_equal(np.array(log), g0["log"])
    assert out.info == int(g0["info"]) and out.tend == float(g0["tend"])
    assert out.nv[1] == g0["nv"][1]
    np.testing.assert_array_equal(out.dtv[2:5], g0["dtv"][2:5])
    np.testing.assert_array_equal(out.cflv[2:4], g0["cflv"][2:4])
    if c["method"][2] < 0:
        np.testing.assert_array_equal(q[2:-2, 2:-2, 2:-2], g0["q"][2:-2, 2:-2, 2:-2])
    assert max_rel_err3(q, g0["q"]) <= 1e-12


def test_user_bc1_and_bc3_equal_builtin():
    """User boundary routines that extrapolate (bc1.f and bc3.f with
    mthbc = 1: x faces, then y over every column, then z) give the built-in
    results bit for bit, retakes included."""
    q0, aux, dx = H.initial1("vc1d_src_strang")

    def bc1(q, aux, t, dt, mthbc, mbc=2):
        q[:mbc] = q[mbc]
        q[-mbc:] = q[-mbc - 1]

    res = []
    for user in (False, True):
        with Grid1D(SOLVER_VC_ACOUSTICS, q0.shape[0] - 4, [], maux=2) as g:
            g.set_aux(aux)
            g.set_q(q0)
            kw = dict(bc1=bc1) if user else {}
            out = g.claw1(dx, 0.0, 0.3, [0.04, 1e99], [1.0, 0.9], 500, (1, 2, 0, 0, 0, 0, 0),
                          (3, 3), (0, 0) if user else (1, 1), **kw)
            res.append((g.get_q()[2:-2], out["nv"][1], out["tend"]))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1:] == res[1][1:]

    def bc3(q, aux, t, dt, mthbc, mbc=2):
        q[:, :, :mbc] = q[:, :, mbc:mbc + 1]
        q[:, :, -mbc:] = q[:, :, -mbc - 1:-mbc]
        q[:, :mbc] = q[:, mbc:mbc + 1]
        q[:, -mbc:] = q[:, -mbc - 1:-mbc]
        q[:mbc] = q[mbc:mbc + 1]
        q[-mbc:] = q[-mbc - 1:-mbc]

    q0, aux = H.initial3("acoustics3d_src_strang_split")
    c = H.CASES3["acoustics3d_src_strang_split"]
    res = []
    for user in (False, True):
        with Grid3D(c["mx"], c["my"], c["mz"], maux=2) as g:
            g.set_aux(aux)
            g.set_q(q0)
            kw = dict(bc3=bc3) if user else {}
            out = g.claw3(0.0, 0.3, [0.3, 1e9], [1.0, 0.9], 500, method=(1, 2, -1, 0, 0, 0, 0),
                          mthlim=(3, 3), mthbc=(0,) * 6 if user else (1,) * 6, **kw)
            res.append((g.get_q()[2:-2, 2:-2, 2:-2], out.nv[1], out.tend))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1:] == res[1][1:]


# ------------------------------------------------------ device-side routines

DEV_CO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "device_hooks.co")


def _dev_hooks(k, ncalls, **names):
    h = gpu.DeviceHooks(DEV_CO, nuser=3 + 3 * ncalls, **names)
    h.write_user([k, 0.0, float(ncalls)])
    return h


def _dev_log(h):
    u = h.read_user()
    n = int(u[1])
    assert n <= int(u[2]), "more calls than the log holds"
    return u[3:3 + 3 * n].reshape(n, 3)


@pytest.mark.parametrize("name", list(H.CASES))
def test_device_hooks_match_reference_claw2(name):
    """The test routines as device kernels in a code object
    (tests/device_hooks.hip; clawgpu_claw2_dev_hooks): the same calls with
    the same (t, dt) as the reference claw2 calling the host routines --
    retakes and Strang half steps included, the pre-source state restored on
    the device for a retake -- and the same final state, bit for bit."""
    c = H.CASES[name]
    g0 = np.load(os.path.join(GOLD, f"claw2_hooks_{name}.npz"))
    q0, dx, dy = H.initial(name)
    euler = c["solver"] == "euler"
    sid, params, mthlim = ((SOLVER_EULER, [1.4, 1.0], [4, 4, 4]) if euler
                           else (SOLVER_ACOUSTICS, [1.0, 4.0], [4, 4]))
    with _dev_hooks(c["k"], len(g0["log"]) + 16, b4step2="test_b4step2", src2="test_src2") as dh, \
            gpu.Grid2D(sid, c["mx"], c["my"], params) as g:
        g.set_q(q0)
        out = g.claw2(dx, dy, 0.0, c["tend"], list(c["dtv"]), [1.0, 0.9], 500, c["method"], mthlim,
                      c["mthbc"], dev_hooks=dh)
        q = g.get_q()
        log = _dev_log(dh)
    print(name, "steps", out["nv"][1], "calls", len(log), flush=True)
    np.testing.assert_array_equal(log, g0["log"])
    assert out["info"] == int(g0["info"]) and out["tend"] == float(g0["tend"])
    assert out["nv"][1] == g0["nv"][1]
    np.testing.assert_array_equal(out["dtv"][2:5], g0["dtv"][2:5])
    np.testing.assert_array_equal(out["cflv"][2:4], g0["cflv"][2:4])
    np.testing.assert_array_equal(interior(q), interior(g0["q"]))


def test_device_bc2_equals_builtin():
    """A device user boundary routine (extrapolation) gives the built-in
    mthbc = 1 result bit for bit, retakes included."""
    q0, dx, dy = H.initial("euler_src_strang")
    method = (1, 2, 2, 0, 0, 0, 0)
    res = []
    for user in (False, True):
        with _dev_hooks(0.0, 1, bc2="test_bc2_extrap") as dh, \
                gpu.Grid2D(SOLVER_EULER, q0.shape[1] - 4, q0.shape[0] - 4, [1.4, 1.0]) as g:
            g.set_q(q0)
            kw = dict(dev_hooks=dh) if user else {}
            out = g.claw2(dx, dy, 0.0, 0.1, [0.02, 1e99], [1.0, 0.9], 500, method, [4, 4, 4],
                          (1, 1, 1, 1) if not user else (0, 0, 0, 0), **kw)
            res.append((interior(g.get_q()), out))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1]["nv"][1] == res[1][1]["nv"][1] and res[0][1]["tend"] == res[1][1]["tend"]


def test_device_hooks_keep_q_on_the_device():
    """At 1024^2 (Euler, src2 every step) the device routines run the loop
    without moving q: bit for bit with the host routines, and faster (each
    host call copies q to the host and back)."""
    import time

    n, k, steps = 1024, 0.9, 12
    from classic_amd import problems as P
    q0, dx, dy = P.euler_quadrants(n, n, seed=5)
    method, mthlim = (1, 2, 2, 0, 1, 0, 0), [4, 4, 4]
    b4, src, log = H.python_hooks(k)
    times, qs = [], []
    for dev in (False, True):
        with _dev_hooks(k, 4 * steps, b4step2="test_b4step2", src2="test_src2") as dh, \
                gpu.Grid2D(SOLVER_EULER, n, n, [1.4, 1.0]) as g:
            g.set_q(q0)
            kw = dict(dev_hooks=dh) if dev else dict(b4step2=b4, src2=src)
            t0 = time.perf_counter()
            out = g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], steps, method, mthlim,
                          (1, 1, 1, 1), **kw)
            times.append(time.perf_counter() - t0)
            qs.append(g.get_q())
            assert out["nv"][1] == steps
    print(f"1024^2, {steps} steps with src2: host routines {times[0] * 1e3:.1f} ms, "
          f"device routines {times[1] * 1e3:.1f} ms", flush=True)
    np.testing.assert_array_equal(interior(qs[0]), interior(qs[1]))
    assert times[1] < times[0]


@pytest.mark.parametrize("name", list(H.CASES3))
def test_device_hooks_match_reference_claw3(name):
    """The 3D test routines as device kernels (clawgpu3_claw3_dev_hooks):
    the same calls with the same (t, dt) as the reference claw3 calling the
    host routines, Strang half steps with the pre-source state kept on the
    device, and the same final state (bit for bit for the split step, to
    1e-12 for the unsplit one, as with host routines)."""
    c = H.CASES3[name]
    g0 = np.load(os.path.join(GOLD, f"claw3_hooks_{name}.npz"))
    q0, aux = H.initial3(name)
    with _dev_hooks(c["k"], len(g0["log"]) + 16, b4step2="test_b4step3", src2="test_src3") as dh, \
            Grid3D(c["mx"], c["my"], c["mz"], maux=2) as g:
        g.set_aux(aux)
        g.set_q(q0)
        out = g.claw3(0.0, c["tend"], list(c["dtv"]), [1.0, 0.9], 500, method=c["method"],
                      mthlim=c["mthlim"], mthbc=c["mthbc"], dev_hooks=dh)
        q = g.get_q()
        log = _dev_log(dh)
    print(name, "steps", out.nv[1], "calls", len(log), flush=True)
    np.testing.assert_array_equal(log, g0["log"])
    assert out.info == int(g0["info"]) and out.tend == float(g0["tend"])
    assert out.nv[1] == g0["nv"][1]
    np.testing.assert_array_equal(out.dtv[2:5], g0["dtv"][2:5])
    np.testing.assert_array_equal(out.cflv[2:4], g0["cflv"][2:4])
    if c["method"][2] < 0:
        np.testing.assert_array_equal(q[2:-2, 2:-2, 2:-2], g0["q"][2:-2, 2:-2, 2:-2])
    assert max_rel_err3(q, g0["q"]) <= 1e-12


def test_device_bc3_equals_builtin():
    """A device bc3 that extrapolates gives the built-in mthbc = 1 result bit
    for bit (split step, retakes included)."""
    q0, aux = H.initial3("acoustics3d_src_strang_split")
    c = H.CASES3["acoustics3d_src_strang_split"]
    res = []
    for user in (False, True):
        with _dev_hooks(0.0, 1, bc2="test_bc3_extrap") as dh, \
                Grid3D(c["mx"], c["my"], c["mz"], maux=2) as g:
            g.set_aux(aux)
            g.set_q(q0)
            kw = dict(dev_hooks=dh) if user else {}
            out = g.claw3(0.0, 0.3, [0.3, 1e9], [1.0, 0.9], 500, method=(1, 2, -1, 0, 0, 0, 0),
                          mthlim=(3, 3), mthbc=(0,) * 6 if user else (1,) * 6, **kw)
            res.append((g.get_q()[2:-2, 2:-2, 2:-2], out.nv[1], out.tend))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1:] == res[1][1:]


def test_device_b4step3_changing_aux_equals_host():
    """A b4step3 that raises every sound speed by (1 + f dt) each attempt:
    as a device kernel (the library recomputes the Courant bound on the
    device and drops the material table once it no longer matches) it gives
    what the same routine gives on the host (aux copied back, table rebuilt),
    bit for bit: the split step, its dt sequence and its retakes."""
    q0, aux = H.initial3("acoustics3d_src_strang_split")
    c = H.CASES3["acoustics3d_src_strang_split"]
    f = 0.25
    calls = []

    def b4(q, a, t, dt):
        calls.append((t, dt))
        a[..., 1] = a[..., 1] * (1.0 + f * dt)

    res = []
    for dev in (False, True):
        with _dev_hooks(f, 400, b4step2="test_b4step3_speed") as dh, \
                Grid3D(c["mx"], c["my"], c["mz"], maux=2) as g:
            g.set_aux(aux)
            g.set_q(q0)
            kw = dict(dev_hooks=dh) if dev else dict(b4step3=b4)
            out = g.claw3(0.0, 0.3, [0.3, 1e9], [1.0, 0.9], 500, method=(1, 2, -1, 0, 0, 0, 0),
                          mthlim=(3, 3), mthbc=(1, 1, 3, 3, 1, 1), **kw)
            log = _dev_log(dh) if dev else np.array([(1.0, t, dt) for t, dt in calls])
            res.append((g.get_q()[2:-2, 2:-2, 2:-2], out, log))
    (qh, oh, lh), (qd, od, ld) = res
    assert len(lh) > oh.nv[1]  # some attempt was retaken
    np.testing.assert_array_equal(ld, lh)
    assert oh.nv[1] == od.nv[1] and oh.tend == od.tend
    np.testing.assert_array_equal(od.dtv[2:5], oh.dtv[2:5])
    np.testing.assert_array_equal(od.cflv[2:4], oh.cflv[2:4])
    np.testing.assert_array_equal(qd, qh)


def test_device_hooks_3d_keep_q_on_the_device():
    """At 192^3 (unsplit step, src3 every step) the device routines run the
    claw3 loop without moving q: the same state as with host routines, bit
    for bit, and faster."""
    import time

    from classic_amd import problems as P
    n, k, steps = 192, 0.6, 8
    q0, aux, _ = P.acoustics3d_pulse(n, seed=3)
    b4, src, log = H.python_hooks(k, dim=3)
    times, qs = [], []
    for dev in (False, True):
        with _dev_hooks(k, 4 * steps, b4step2="test_b4step3", src2="test_src3") as dh, \
                Grid3D(n, n, n, maux=2) as g:
            g.set_aux(aux)
            g.set_q(q0)
            kw = dict(dev_hooks=dh) if dev else dict(b4step3=b4, src3=src)
            t0 = time.perf_counter()
            out = g.claw3(0.0, 1e9, [1e-3, 1e99], [1.0, 0.9], steps, method=(1, 2, 22, 0, 1, 0, 2),
                          mthlim=(2, 2), mthbc=(3,) * 6, **kw)
            times.append(time.perf_counter() - t0)
            qs.append(g.get_q())
            assert out.nv[1] == steps
    print(f"192^3, {steps} steps with src3: host routines {times[0] * 1e3:.1f} ms, "
          f"device routines {times[1] * 1e3:.1f} ms", flush=True)
    np.testing.assert_array_equal(qs[0][2:-2, 2:-2, 2:-2], qs[1][2:-2, 2:-2, 2:-2])
    assert times[1] < times[0]


def test_device_hooks_load_errors():
    """A missing code object or kernel is an error of the load call alone:
    the next step on the device runs normally."""
    from classic_amd._lib import ClawGpuError
    with pytest.raises(ClawGpuError):
        gpu.DeviceHooks(DEV_CO, src2="no_such_kernel")
    with pytest.raises(ClawGpuError):
        gpu.DeviceHooks(DEV_CO + ".missing", src2="test_src2")
    q0, dx, dy = H.initial("euler_src_strang")
    with gpu.Grid2D(SOLVER_EULER, q0.shape[1] - 4, q0.shape[0] - 4, [1.4, 1.0]) as g:
        g.set_q(q0)
        out = g.claw2(dx, dy, 0.0, 0.02, [0.02, 1e99], [1.0, 0.9], 10, (1, 2, 2, 0, 0, 0, 0),
                      [4, 4, 4], (1, 1, 1, 1))
        assert out["tend"] == pytest.approx(0.02) and np.isfinite(g.get_q()).all()
