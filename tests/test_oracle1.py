"""CPU checks of the 1D oracle (oracle/claw1d_oracle.c): the two acoustics
examples' regression frame sums (copied into tests/golden), bit-for-bit
agreement with the golden vectors the reference Fortran wrote
(tests/golden/make_golden1.py) and, when oracle/_ref is built, with the live
reference; closed forms of the rp1 solvers and of the upwind scheme."""
import os

import numpy as np
import pytest

from classic_amd import problems as P
from oracle import oracle1 as O1

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def test_acoustics_example1_frames_and_regression_sums():
    g = gold("acoustics1d_example1")
    q, dx = P.acoustics1d_gaussian(400)
    p = O1.Problem1(400, O1.SOLVER_ACOUSTICS, O1.params1(1.0, 4.0), mthlim=(4, 4))
    fr = O1.frames(p, q, dx, 0.0, 1.0, 2, 0.1, nmax=500)
    for k, (t, qq) in enumerate(fr):
        assert t == g["t"][k]
        np.testing.assert_array_equal(qq[2:-2], g["q"][k][2:-2])
        sums = qq[2:-2].sum(axis=0)
        reg = g["regression_sums"][k]
        np.testing.assert_allclose(sums, reg, rtol=1e-14, atol=1e-14)


def test_acoustics_heterogeneous_frames_and_regression_sums():
    g = gold("acoustics1d_heterogeneous")
    q, aux, dx = P.acoustics1d_heterogeneous(500)
    p = O1.Problem1(500, O1.SOLVER_VC_ACOUSTICS, maux=2, mthlim=(4, 4))
    fr = O1.frames(p, q, dx, 0.0, 5.0, 2, 1.0, dt_max=1e9, aux=aux)
    for k, (t, qq) in enumerate(fr):
        assert t == g["t"][k]
        np.testing.assert_array_equal(qq[2:-2], g["q"][k][2:-2])
        np.testing.assert_allclose(qq[2:-2].sum(axis=0), g["regression_sums"][k], rtol=1e-14,
                                   atol=1e-14)


@pytest.mark.parametrize("case", range(6))
def test_golden_step1(case):
    d = gold(f"step1_case{case}")
    sid = int(d["solver"])
    mw = 1 if sid == O1.SOLVER_ADVECTION else 2
    aux = d["aux"] if d["aux"].size else None
    maux = aux.shape[1] if aux is not None else 0
    p = O1.Problem1(57, sid, O1.params1(1.3, 2.9, u=-0.7), meqn=mw, mwaves=mw, maux=maux,
                    method=tuple(int(x) for x in d["method"]),
                    mthlim=tuple(int(x) for x in d["mthlim"]), use_fwave=bool(d["use_fwave"]))
    qo, cfl = O1.step1(p, d["q_in"], 2.0 / 57, 0.01, aux=aux)
    np.testing.assert_array_equal(qo, d["q_out"])
    assert cfl == d["cfl"]


@pytest.mark.skipif(not O1.ref_available(), reason="reference build oracle/_ref absent")
def test_live_reference_claw1_with_retakes():
    rng = np.random.default_rng(4)
    q = rng.standard_normal((104, 2))
    aux = np.stack([rng.uniform(0.5, 3, 104), rng.uniform(0.5, 2, 104)], -1)
    p = O1.Problem1(100, O1.SOLVER_VC_ACOUSTICS, maux=2, mthlim=(2, 3), mthbc=(3, 2))
    p.mthbc = O1._arr_i((3, 3))
    a = O1.claw1(p, q, 0.01, 0.0, 0.3, [0.5, 1e9], [1.0, 0.9], 1000, aux=aux)
    b = O1.claw1(p, q, 0.01, 0.0, 0.3, [0.5, 1e9], [1.0, 0.9], 1000, aux=aux, backend="ref")
    assert a.info == b.info == 0 and a.tend == b.tend
    np.testing.assert_array_equal(a.q[2:-2], b.q[2:-2])
    np.testing.assert_array_equal(a.dtv, b.dtv)
    np.testing.assert_array_equal(a.cflv, b.cflv)
    np.testing.assert_array_equal(a.nv, b.nv)
    assert len(a.dt_trace) > a.nv[1]  # the run had retakes


@pytest.mark.parametrize("mthbc", [(1, 1), (2, 2), (3, 3), (1, 3)])
def test_bc1_matches_reference_or_closed_form(mthbc):
    rng = np.random.default_rng(5)
    p = O1.Problem1(9, mthbc=mthbc)
    q = rng.standard_normal((13, 2))
    out = O1.bc1(p, q)
    i = np.arange(13) - 1  # Fortran index of row r is r - 1
    for r in (0, 1, 11, 12):
        lo = r < 2
        code = mthbc[0 if lo else 1]
        src = {1: 2 if lo else 10, 2: r + 9 if lo else r - 9, 3: 3 - r if lo else 21 - r}[code]
        want = q[src].copy()
        if code == 3:
            want[1] = -want[1]
        np.testing.assert_array_equal(out[r], want, err_msg=f"row {r} (i={i[r]})")
    if O1.ref_available():
        np.testing.assert_array_equal(out, O1.bc1(p, q, backend="ref"))


def test_rp1_identities():
    rng = np.random.default_rng(6)
    q = rng.standard_normal((20, 2))
    aux = np.stack([rng.uniform(0.5, 3, 20), rng.uniform(0.5, 2, 20)], -1)
    wave, s, amdq, apdq = O1.rp1(O1.SOLVER_VC_ACOUSTICS, O1.params1(), q, aux)
    dq = q[1:] - q[:-1]
    np.testing.assert_allclose(wave[1:].sum(axis=1), dq, rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(s[1:, 0], -aux[:-1, 1])
    np.testing.assert_array_equal(s[1:, 1], aux[1:, 1])
    # constant coefficients: amdq + apdq = A dq with A = [[0, K], [1/rho, 0]]
    rho, K = 1.7, 3.1
    wave, s, amdq, apdq = O1.rp1(O1.SOLVER_ACOUSTICS, O1.params1(rho, K), q)
    adq = np.stack([K * dq[:, 1], dq[:, 0] / rho], -1)
    np.testing.assert_allclose((amdq + apdq)[1:], adq, rtol=1e-13, atol=1e-14)
    # advection: the single wave is the jump, fluctuations split by the sign of u
    wave, s, amdq, apdq = O1.rp1(O1.SOLVER_ADVECTION, O1.params1(u=-0.4), q[:, :1], mwaves=1)
    np.testing.assert_array_equal(wave[1:, 0, 0], q[1:, 0] - q[:-1, 0])
    np.testing.assert_array_equal(apdq[1:, 0], 0.0 * wave[1:, 0, 0])
    np.testing.assert_allclose(amdq[1:, 0], -0.4 * wave[1:, 0, 0], rtol=1e-15)


def test_upwind_at_courant_one_is_an_exact_shift():
    rng = np.random.default_rng(7)
    q = np.zeros((44, 1))
    q[2:-2, 0] = rng.standard_normal(40)
    p = O1.Problem1(40, O1.SOLVER_ADVECTION, O1.params1(u=1.0), meqn=1, mwaves=1,
                    method=(1, 1, 0, 0, 0, 0, 0), mthlim=(0,), mthbc=(2, 2))
    q = O1.bc1(p, q)
    out, cfl = O1.step1(p, q, 0.1, 0.1)
    assert cfl == 1.0
    # q_i - (q_i - q_{i-1}) rounds once: equal to q_{i-1} within an ulp of |q|
    np.testing.assert_allclose(out[3:-2, 0], q[2:-3, 0], rtol=0, atol=4e-16 * np.abs(q).max())
