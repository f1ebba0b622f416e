"""User routines of the plug-in contract (bc2, b4step2, src2) on the device
claw2 loop (clawgpu_claw2_hooks) against the reference claw2 run with the
same routines (tests/golden/claw2_hooks_*.npz, written by the reference
build through tests/golden/make_golden_hooks.py): the same calls in the same
order with the same (t, dt) — retakes and Strang half steps included — and
the same final state bit for bit (EXACT arithmetic)."""
import os

import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from tests import hooks_cases as H
from tests.helpers import interior

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
