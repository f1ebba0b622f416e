"""CPU checks of the 3D oracle (oracle/claw3d_oracle.c).

1. It reproduces the regression frame sums that the reference example holds
   (examples/acoustics_3d_heterogeneous/regression_data, copied into
   tests/golden/hetero3d_frames_20.npz) -- the pin for the restated
   variable-coefficient rpn3/rpt3/rptt3, whose upstream source is absent.
2. It reproduces, bit for bit, the golden vectors the reference Fortran
   (oracle/_ref, built from /root/reference) wrote into tests/golden
   (tests/golden/make_golden3.py), and agrees live with that build when it is
   present.
3. Closed forms of the solvers: waves sum to the jump, amdq + apdq = A dq and
   bm + bp = B asdq in a homogeneous medium, transmission at an interface.
4. Properties: conservation with periodic boundaries in a homogeneous medium,
   invariance under a cyclic rotation of the axes.
"""
import os

import numpy as np
import pytest

from classic_amd import problems as P
from oracle import oracle3 as O3

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def problem_of(d, mx, my, mz, mthbc=(1, 1, 1, 1, 1, 1)):
    m = [int(x) for x in d["method"]]
    maux = d["aux"].shape[-1]
    return O3.Problem3(mx, my, mz, maux=maux, mthlim=tuple(int(x) for x in d["mthlim"]),
                       method=tuple(m), use_fwave=bool(d["use_fwave"]), mthbc=mthbc)


# --------------------------------------------------------------- regression

def test_regression_frame_sums_of_the_heterogeneous_example():
    g = gold("hetero3d_frames_20")
    p = O3.Problem3(20, 20, 20, mthbc=(1, 1, 1, 1, 1, 0), mthlim=(3, 3))
    q, aux = P.acoustics3d_heterogeneous(20)
    fr = O3.frames(p, q, aux, 0.0, 1.0, 2, 0.01)
    for k, (t, qq) in enumerate(fr):
        assert t == g["t"][k]
        sums = np.array([qq[2:-2, 2:-2, 2:-2, m].sum() for m in range(4)])
        np.testing.assert_array_equal(sums, g["sums"][k])          # reference build, bitwise
        reg = g["regression_sums"][k]                              # the example's own file
        assert abs(sums[0] - reg[0]) <= 1e-14 * abs(reg[0])
        assert np.all(np.abs(sums[1:3] - reg[1:3]) <= 1e-15)
    np.testing.assert_array_equal(fr[0][1][2:-2:3, 2:-2:3, 2:-2:3], g["q_frame1_sub"])


# -------------------------------------------------------------- golden steps

@pytest.mark.parametrize("case", range(10))
def test_golden_step3_and_dimsp3(case):
    d = gold(f"step3_case{case}")
    p = problem_of(d, 6, 5, 4)
    if p.method[2] >= 0:
        qo, cfl = O3.step3(p, d["q_in"], d["aux"], 0.02)
    else:
        qo, cfl = O3.dimsp3(p, d["q_in"], d["aux"], 0.02)
    np.testing.assert_array_equal(qo, d["q_out"])
    assert cfl == d["cfl"]


@pytest.mark.parametrize("case", range(4))
def test_golden_bc3(case):
    d = gold(f"bc3_case{case}")
    p = O3.Problem3(5, 4, 6, mthbc=tuple(int(x) for x in d["mthbc"]), lower=(-0.2, -0.1, -1.0))
    np.testing.assert_array_equal(O3.bc3(p, d["q_in"], float(d["t"])), d["q_out"])


def test_golden_claw3_with_retakes():
    d = gold("claw3_pulse_12x10x8")
    p = O3.Problem3(12, 10, 8, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2))
    r = O3.claw3(p, d["q_in"], d["aux"], 0.0, 0.5, [0.2, 1e9], [1.0, 0.9], 500)
    assert r.info == d["info"] and r.tend == d["tend"]
    np.testing.assert_array_equal(r.q, d["q_out"])
    np.testing.assert_array_equal(r.dtv, d["dtv"])
    np.testing.assert_array_equal(r.cflv, d["cflv"])
    np.testing.assert_array_equal(r.nv, d["nv"])
    assert len(r.dt_trace) > r.nv[1]  # the run had retakes


@pytest.mark.skipif(not O3.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("seed", [5, 6])
def test_live_reference_agreement(seed):
    rng = np.random.default_rng(seed)
    for meth, lim in [((1, 2, 22), (3, 4)), ((1, 2, 21), (2, 2)), ((1, 2, -1), (1, 3))]:
        p = O3.Problem3(5, 7, 6, mthlim=lim, method=meth + (0, 0, 0, 0))
        q = rng.standard_normal(p.shape())
        aux = np.stack([rng.uniform(0.5, 3, p.shape()[:3]), rng.uniform(0.5, 2, p.shape()[:3])], -1)
        f = O3.step3 if meth[2] >= 0 else O3.dimsp3
        a, ca = f(p, q, aux, 0.03)
        b, cb = f(p, q, aux, 0.03, backend="ref")
        np.testing.assert_array_equal(a, b)
        assert ca == cb


# ------------------------------------------------------------ solver identities

@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_rpn3_waves_sum_to_jump_and_speeds(ixyz):
    rng = np.random.default_rng(ixyz)
    n = 12
    q = rng.standard_normal((n, 4))
    aux = np.stack([rng.uniform(0.5, 3, n), rng.uniform(0.5, 2, n)], -1)
    wave, s, amdq, apdq = O3.rpn3(ixyz, q, aux)
    mu = ixyz  # 0-based normal velocity component
    for k in range(1, n):
        dq = q[k] - q[k - 1]
        tot = wave[k].sum(axis=0)
        assert abs(tot[0] - dq[0]) <= 1e-14 * (1 + abs(dq[0]))
        assert abs(tot[mu] - dq[mu]) <= 1e-14 * (1 + abs(dq[mu]))
        # left-going wave is an eigenvector (-Z_{k-1}, 1) of cell k-1, right-going (Z_k, 1)
        assert np.isclose(wave[k, 0, 0], -aux[k - 1, 0] * wave[k, 0, mu], rtol=1e-14, atol=0)
        assert np.isclose(wave[k, 1, 0], aux[k, 0] * wave[k, 1, mu], rtol=1e-14, atol=0)
        assert s[k, 0] == -aux[k - 1, 1] and s[k, 1] == aux[k, 1]


@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_homogeneous_medium_flux_difference_is_A_dq(ixyz):
    rng = np.random.default_rng(10 + ixyz)
    n, Z, c = 10, 1.7, 1.3
    q = rng.standard_normal((n, 4))
    aux = np.tile([Z, c], (n, 1))
    _, _, amdq, apdq = O3.rpn3(ixyz, q, aux)
    mu = ixyz
    for k in range(1, n):
        dq = q[k] - q[k - 1]
        # p_t + K u_x = 0, u_t + p_x / rho = 0 with K = Z c, 1/rho = c / Z
        adq = np.zeros(4)
        adq[0] = Z * c * dq[mu]
        adq[mu] = c / Z * dq[0]
        np.testing.assert_allclose(amdq[k] + apdq[k], adq, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("ixyz,icoor", [(1, 2), (1, 3), (2, 2), (2, 3), (3, 2), (3, 3)])
def test_homogeneous_transverse_split_is_B_asdq(ixyz, icoor):
    rng = np.random.default_rng(20 + 3 * ixyz + icoor)
    n, Z, c = 8, 2.5, 0.7
    asdq = rng.standard_normal((n, 4))
    comps = {1: (1, 2, 3), 2: (2, 3, 1), 3: (3, 1, 2)}[ixyz]  # (mu, mv, mw), 0-based
    mt = comps[1] if icoor == 2 else comps[2]
    asdq[:, comps[0]] = 0.0
    asdq[:, comps[2] if icoor == 2 else comps[1]] = 0.0
    aux = np.tile([Z, c], (3, 3, n, 1))
    for imp in (1, 2):
        bm, bp = O3.rpt3(ixyz, icoor, imp, aux, asdq)
        b = np.zeros((n, 4))
        b[:, 0] = Z * c * asdq[:, mt]
        b[:, mt] = c / Z * asdq[:, 0]
        np.testing.assert_allclose((bm + bp)[1:], b[1:], rtol=1e-13, atol=1e-14)
        # down-going part moves at -c, up-going at +c: bm = -c * r1, bp = c * r2
        np.testing.assert_allclose(bm[1:, 0], -Z * bm[1:, mt], rtol=1e-13)
        np.testing.assert_allclose(bp[1:, 0], Z * bp[1:, mt], rtol=1e-13)


def test_transverse_split_uses_the_cell_the_fluctuation_moved_into():
    """imp=1 (amdq) reads cell i-1's column of materials, imp=2 cell i's;
    rptt3 with icoor=2 reads the neighbour plane chosen by impt."""
    n = 6
    asdq = np.zeros((n, 4))
    asdq[:, 0] = 1.0
    aux = np.ones((3, 3, n, 2))
    aux[0, 1, 2, :] = (4.0, 3.0)  # aux1 (row below), middle plane, cell 2
    # array row r is cell r-1 (mbc = 2): row 2 = cell 1
    bm1, _ = O3.rpt3(1, 2, 1, aux, asdq)   # interface of row 3 reads cell 1 for imp=1
    bm2, _ = O3.rpt3(1, 2, 2, aux, asdq)   # interface of row 2 reads cell 1 for imp=2
    # a1 = -1/(zm+zz); bm(p) = cm*a1*zm
    assert bm1[3, 0] == 3.0 * (-1.0 / (4.0 + 1.0)) * 4.0
    assert bm2[2, 0] == 3.0 * (-1.0 / (4.0 + 1.0)) * 4.0
    assert bm1[2, 0] == -0.5 and bm2[3, 0] == -0.5
    aux = np.ones((3, 3, n, 2))
    aux[0, 2, 2, :] = (4.0, 3.0)  # aux1, plane k+1, cell 2
    bmt, _ = O3.rpt3(1, 2, 2, aux, asdq, impt=2)
    bmt1, _ = O3.rpt3(1, 2, 2, aux, asdq, impt=1)
    assert bmt[2, 0] == 3.0 * (-1.0 / 5.0) * 4.0
    assert bmt1[2, 0] == -0.5


# ----------------------------------------------------------------- properties

def test_periodic_homogeneous_conservation():
    rng = np.random.default_rng(3)
    p = O3.Problem3(8, 7, 6, mthbc=(2,) * 6, mthlim=(4, 4))
    q = O3.bc3(p, rng.standard_normal(p.shape()))
    aux = np.tile([1.5, 1.2], p.shape()[:3] + (1,))
    qo, _ = O3.step3(p, q, aux, 0.05)
    s0 = q[2:-2, 2:-2, 2:-2].sum(axis=(0, 1, 2))
    s1 = qo[2:-2, 2:-2, 2:-2].sum(axis=(0, 1, 2))
    np.testing.assert_allclose(s1, s0, rtol=0, atol=1e-12 * np.abs(q).sum())


def test_cyclic_axis_rotation_invariance():
    """(x, y, z) -> (y, z, x) with (u, v, w) -> (v, w, u) maps the step onto
    itself up to summation order."""
    rng = np.random.default_rng(4)
    mx, my, mz = 6, 5, 4
    p = O3.Problem3(mx, my, mz, mthlim=(3, 3))
    q = O3.bc3(p, rng.standard_normal(p.shape()))
    aux = np.stack([rng.uniform(0.5, 3, p.shape()[:3]), rng.uniform(0.5, 2, p.shape()[:3])], -1)
    qo, cfl = O3.step3(p, q, aux, 0.03)
    # array axes are (z, y, x): new x' = old y, y' = old z, z' = old x
    pr = O3.Problem3(my, mz, mx, mthlim=(3, 3))
    rot = lambda a: np.ascontiguousarray(np.transpose(a, (2, 0, 1, 3)))  # noqa: E731
    qr = rot(q)[..., [0, 2, 3, 1]]
    qro, cflr = O3.step3(pr, qr, rot(aux), 0.03)
    back = rot(qo)[..., [0, 2, 3, 1]]
    assert abs(cfl - cflr) <= 1e-15
    np.testing.assert_allclose(qro[2:-2, 2:-2, 2:-2], back[2:-2, 2:-2, 2:-2], rtol=0,
                               atol=1e-13 * np.abs(q).max())
