"""CPU pinning of the oracle (no GPU).

1. The oracle reproduces, bit for bit, the golden vectors that the reference
   Fortran (oracle/_ref, built from /root/reference) wrote into tests/golden
   (tests/golden/make_golden.py), and agrees live with that build when it is
   present.
2. Facts that fix the arithmetic independently of any implementation:
   closed-form limiter values, TVD bounds and symmetry; Riemann-solver
   identities (sum of waves = jump, amdq + apdq = A dq or f(qr) - f(ql),
   transverse split = B asdq, rotation invariance); the exact solution of the
   Euler Riemann problem (Toro test 1, Sod) and of a travelling acoustic
   plane wave, which the scheme converges to.
"""
import os

import numpy as np
import pytest

from classic_amd import problems as P
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 1.4


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def acoustics_problem(mx, my, method=(1, 2, 2, 0, 0, 0, 0), lim=(4, 4), mthbc=(3, 3, 3, 3),
                      rho=1.0, bulk=4.0):
    return O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(rho, bulk), method=method,
                     mthlim=list(lim), mthbc=mthbc)


def euler_problem(mx, my, method=(1, 2, 2, 0, 0, 0, 0), lim=(4, 4, 4), mthbc=(1, 1, 1, 1),
                  efix=True):
    return O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=G, efix=efix),
                     method=method, mthlim=list(lim), mthbc=mthbc)


# ------------------------------------------------------------ golden vectors

def _check_claw(g, p, nmax, tend, dtv0):
    r = O.claw2(p, g["q_in"], float(g["dx"]), float(g["dy"]), 0.0, tend, dtv0, [1.0, 0.9], nmax)
    np.testing.assert_array_equal(r.q[2:-2, 2:-2], g["q_out"])
    assert r.info == int(g["info"])
    np.testing.assert_array_equal(r.dtv, g["dtv"])
    np.testing.assert_array_equal(r.cflv, g["cflv"])
    np.testing.assert_array_equal(r.nv, g["nv"])


def test_golden_acoustics_pulse_40():
    _check_claw(gold("acoustics_pulse_40"), acoustics_problem(40, 40), 50000, 0.3, [1e-3, 1e99])


def test_golden_config1_acoustics_200():
    g = gold("config1_acoustics_200")
    q, dx, dy = P.acoustics_pulse(200, 200)
    r = O.claw2(acoustics_problem(200, 200), q, dx, dy, 0.0, 0.6, [1e-3, 1e99], [1.0, 0.9], 50000)
    qo = r.q[2:-2, 2:-2]
    np.testing.assert_array_equal(qo.sum(axis=(0, 1)), g["sums"])
    np.testing.assert_array_equal(qo[::10, ::10], g["sub"])
    np.testing.assert_array_equal(r.dtv, g["dtv"])
    assert int(r.nv[1]) == int(g["nv"][1]) == 135
    assert r.info == 0


def test_golden_euler_quadrants():
    _check_claw(gold("euler_quadrants_32x24"), euler_problem(32, 24), 20, 1e9, [1e-4, 1e99])


@pytest.mark.parametrize("split", [1, 2])
def test_golden_dimensional_splitting(split):
    _check_claw(gold(f"euler_dimsplit{split}_24x20"),
                euler_problem(24, 20, method=(1, 2, -split, 0, 0, 0, 0)), 10, 1e9, [1e-4, 1e99])


def test_golden_step2_all_methods_and_limiters():
    g = gold("euler_step2_17x13")
    qb = g["q_in"]
    dx, dy = float(g["dx"]), float(g["dy"])
    for meth in [(1, 2, 2), (1, 2, 1), (1, 2, 0), (1, 1, 1), (1, 1, 0)]:
        for lim in range(6):
            p = euler_problem(17, 13, method=meth + (0, 0, 0, 0), lim=(lim,) * 3,
                              mthbc=(3, 3, 2, 2))
            qn, cfl = O.step2(p, qb, dx, dy, float(g["dt"]))
            key = f"m{meth[1]}{meth[2]}_l{lim}"
            np.testing.assert_array_equal(qn[2:-2, 2:-2], g["q_" + key], err_msg=key)
            assert cfl == float(g["cfl_" + key])


def test_golden_bc2():
    g = gold("bc2_11x13")
    for mthbc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2), (3, 1, 3, 1)]:
        p = acoustics_problem(11, 13, mthbc=mthbc)
        np.testing.assert_array_equal(O.bc2(p, g["q_in"]), g["q_" + "".join(map(str, mthbc))])


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("seed,mthbc", [(6, (1, 3, 2, 2)), (7, (1, 1, 1, 1)), (3, (3, 3, 3, 3)),
                                        (4, (2, 2, 3, 3))])
def test_live_reference_agreement(seed, mthbc):
    """Against the reference build itself.  The configurations keep the Roe
    solver in physical states: the reference spins forever in its retake
    loop if a NaN appears, so a run that the oracle reports as non-finite
    (info = -5) is not sent to it."""
    q, dx, dy = P.euler_quadrants(45, 38, seed=seed)
    for meth, lim in [((1, 2, 2, 0, 0, 0, 0), (4, 4, 4)), ((1, 2, 1, 0, 0, 0, 0), (2, 1, 3)),
                      ((1, 2, -2, 0, 0, 0, 0), (4, 4, 4))]:
        p = euler_problem(45, 38, method=meth, lim=lim, mthbc=mthbc)
        a = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 15)
        assert a.info == 11 and not np.isnan(a.q).any()
        b = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 15, backend="ref")
        np.testing.assert_array_equal(a.q[2:-2, 2:-2], b.q[2:-2, 2:-2])
        np.testing.assert_array_equal(a.dtv, b.dtv)


# ------------------------------------------------------------------- limiter

def _phi(lim, r):
    """Limiter through the oracle: three unit waves along one direction with
    the upwind (left) neighbour r times the middle wave."""
    n = 5  # cells -> interfaces 1..4 with mbc=2 means mx=1
    wave = np.zeros((n, 1, 1))
    wave[1, 0, 0] = r  # interface 0 (upwind of interface 1)
    wave[2, 0, 0] = 1.0  # interface 1
    wave[3, 0, 0] = 0.0
    s = np.ones((n, 1))
    w = O.limiter(wave, s, [lim])
    return w[2, 0, 0]


R = [-2.0, -0.5, 0.0, 0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 10.0]
CLOSED = {
    1: lambda r: max(0.0, min(1.0, r)),
    2: lambda r: max(0.0, min(1.0, 2 * r), min(2.0, r)),
    3: lambda r: (r + abs(r)) / (1 + abs(r)),
    4: lambda r: max(0.0, min((1 + r) / 2, 2.0, 2 * r)),
    5: lambda r: r,
}
TABLE = {  # phi at R, worked by hand from the definitions
    1: [0, 0, 0, 0.25, 0.5, 1, 1, 1, 1, 1],
    2: [0, 0, 0, 0.5, 1, 1, 1.5, 2, 2, 2],
    3: [0, 0, 0, 0.4, 2 / 3, 1, 1.2, 4 / 3, 1.5, 20 / 11],
    4: [0, 0, 0, 0.5, 0.75, 1, 1.25, 1.5, 2, 2],
    5: R,
}


@pytest.mark.parametrize("lim", [1, 2, 3, 4, 5])
def test_limiter_known_answers(lim):
    got = [_phi(lim, r) for r in R]
    np.testing.assert_allclose(got, TABLE[lim], rtol=1e-15, atol=0)
    np.testing.assert_allclose(got, [CLOSED[lim](r) for r in R], rtol=1e-15, atol=0)


@pytest.mark.parametrize("lim", [1, 2, 3, 4])
def test_limiters_tvd_and_symmetric(lim):
    for r in np.linspace(-3, 6, 91):
        phi = _phi(lim, r)
        assert 0.0 <= phi <= max(0.0, min(2.0 * r, 2.0)) + 1e-15  # Sweby region
        if r > 0:
            assert abs(phi / r - _phi(lim, 1.0 / r)) <= 1e-14  # phi(r)/r = phi(1/r)


def test_limiter_uses_downwind_neighbour_for_negative_speed():
    wave = np.zeros((5, 1, 1))
    wave[1, 0, 0], wave[2, 0, 0], wave[3, 0, 0] = 7.0, 1.0, 0.5
    s = -np.ones((5, 1))
    w = O.limiter(wave, s, [1])  # minmod(r) with r = 0.5 from the right
    assert w[2, 0, 0] == 0.5


# --------------------------------------------------------- Riemann solvers

def cons(rho, u, v, p):
    return np.stack([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)], -1)


def euler_flux(q, mu):
    rho, mx_, my_, E = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    p = (G - 1) * (E - 0.5 * (mx_ ** 2 + my_ ** 2) / rho)
    un = q[..., mu] / rho
    f = q * un[..., None]
    f[..., mu] += p
    f[..., 3] += p * un
    return f


@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("efix", [True, False])
def test_euler_roe_identities(ixy, efix):
    rng = np.random.default_rng(3)
    n = 200
    q = cons(rng.uniform(0.1, 2, n), rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n),
             rng.uniform(0.1, 2, n))
    par = O.solver_params(gamma=G, efix=efix)
    w, s, am, ap = O.rpn2(O.SOLVER_EULER, par, ixy, q, 3)
    dq = q[1:] - q[:-1]
    np.testing.assert_allclose(w[1:].sum(axis=1), dq, atol=1e-13)  # sum of waves = jump
    mu = 1 if ixy == 1 else 2
    df = euler_flux(q[1:], mu) - euler_flux(q[:-1], mu)
    np.testing.assert_allclose(am[1:] + ap[1:], df, atol=1e-12)  # Roe property (efix keeps it)
    # speeds ordered u-a <= u <= u+a, and each wave is an eigenvector: f'(W) ~ s W
    assert np.all(s[1:, 0] <= s[1:, 1]) and np.all(s[1:, 1] <= s[1:, 2])


def test_euler_rotation_invariance():
    rng = np.random.default_rng(4)
    n = 50
    q = cons(rng.uniform(0.2, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
             rng.uniform(0.2, 2, n))
    qr = q[:, [0, 2, 1, 3]]
    par = O.solver_params(gamma=G, efix=True)
    w1, s1, a1, p1 = O.rpn2(O.SOLVER_EULER, par, 1, q, 3)
    w2, s2, a2, p2 = O.rpn2(O.SOLVER_EULER, par, 2, qr, 3)
    np.testing.assert_allclose(s1, s2, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(a1, a2[:, [0, 2, 1, 3]], rtol=1e-13, atol=1e-14)
    asdq = rng.standard_normal((n, 4))
    b1 = O.rpt2(O.SOLVER_EULER, par, 1, 1, q, asdq, 3)
    b2 = O.rpt2(O.SOLVER_EULER, par, 2, 1, qr, asdq[:, [0, 2, 1, 3]], 3)
    np.testing.assert_allclose(b1[0], b2[0][:, [0, 2, 1, 3]], rtol=1e-13, atol=1e-14)


def test_euler_transverse_split_is_B_times_asdq():
    """bmasdq + bpasdq = B(roe) asdq with B the y-flux Jacobian at the Roe
    state of the interface, and bm (bp) uses only the negative (positive)
    eigenvalues: B- + B+ = B."""
    rng = np.random.default_rng(9)
    n = 40
    q = cons(rng.uniform(0.3, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
             rng.uniform(0.3, 2, n))
    asdq = rng.standard_normal((n, 4))
    bm, bp = O.rpt2(O.SOLVER_EULER, O.solver_params(gamma=G), 1, 1, q, asdq, 3)
    for k in range(1, n):
        ql, qr = q[k - 1], q[k]
        sl, sr = np.sqrt(ql[0]), np.sqrt(qr[0])
        u = (ql[1] / sl + qr[1] / sr) / (sl + sr)
        v = (ql[2] / sl + qr[2] / sr) / (sl + sr)
        pl = (G - 1) * (ql[3] - 0.5 * (ql[1] ** 2 + ql[2] ** 2) / ql[0])
        pr = (G - 1) * (qr[3] - 0.5 * (qr[1] ** 2 + qr[2] ** 2) / qr[0])
        H = ((ql[3] + pl) / sl + (qr[3] + pr) / sr) / (sl + sr)
        g1 = G - 1
        k2 = 0.5 * (u * u + v * v)
        B = np.array([[0, 0, 1, 0],
                      [-u * v, v, u, 0],
                      [g1 * k2 - v * v, -g1 * u, (3 - G) * v, g1],
                      [v * (g1 * k2 - H), -g1 * u * v, H - g1 * v * v, G * v]])
        np.testing.assert_allclose(bm[k] + bp[k], B @ asdq[k], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ixy", [1, 2])
def test_acoustics_closed_form(ixy):
    rho, bulk = 2.0, 5.0
    c, Z = np.sqrt(bulk / rho), rho * np.sqrt(bulk / rho)
    rng = np.random.default_rng(2)
    q = rng.standard_normal((60, 3))
    par = O.solver_params(rho, bulk)
    w, s, am, ap = O.rpn2(O.SOLVER_ACOUSTICS, par, ixy, q, 2)
    mu = 1 if ixy == 1 else 2
    A = np.zeros((3, 3))
    A[0, mu], A[mu, 0] = bulk, 1.0 / rho
    dq = q[1:] - q[:-1]
    mv = 3 - mu
    # the two acoustic waves carry the jump in p and the normal velocity; the
    # transverse velocity jump travels with the (unrepresented) zero-speed wave
    np.testing.assert_allclose(w[1:, :, [0, mu]].sum(axis=1), dq[:, [0, mu]], atol=1e-14)
    assert np.all(w[1:, :, mv] == 0.0)
    np.testing.assert_allclose(am[1:] + ap[1:], dq @ A.T, atol=1e-13)
    np.testing.assert_allclose(s[1:], np.tile([-c, c], (59, 1)), rtol=1e-15)
    Bm = np.zeros((3, 3))
    Bm[0, mv], Bm[mv, 0] = bulk, 1.0 / rho
    asdq = rng.standard_normal((60, 3))
    bm, bp = O.rpt2(O.SOLVER_ACOUSTICS, par, ixy, 1, q, asdq, 2)
    np.testing.assert_allclose(bm[2:-2] + bp[2:-2], asdq[2:-2] @ Bm.T, atol=1e-13)
    assert Z > 0


# ---------------------------------------------- exact solutions (convergence)

def exact_euler_riemann(rl, ul, pl, rr, ur, pr, xi):
    """Exact solution of the 1D Euler Riemann problem at x/t = xi (Toro,
    'Riemann Solvers and Numerical Methods for Fluid Dynamics', ch. 4)."""
    cl, cr = np.sqrt(G * pl / rl), np.sqrt(G * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((G + 1) * rk), (G - 1) / (G + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (G - 1) * ((p / pk) ** ((G - 1) / (2 * G)) - 1)

    lo, hi = 1e-10, 100.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if f(pm, rl, pl, cl) + f(pm, rr, pr, cr) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    pm = 0.5 * (lo + hi)
    um = 0.5 * (ul + ur) + 0.5 * (f(pm, rr, pr, cr) - f(pm, rl, pl, cl))
    out = np.empty((len(xi), 3))
    for k, x in enumerate(xi):
        if x < um:  # left of contact
            if pm > pl:
                sl = ul - cl * np.sqrt((G + 1) / (2 * G) * pm / pl + (G - 1) / (2 * G))
                rho = rl * (pm / pl + (G - 1) / (G + 1)) / ((G - 1) / (G + 1) * pm / pl + 1)
                out[k] = (rl, ul, pl) if x < sl else (rho, um, pm)
            else:
                cml = cl * (pm / pl) ** ((G - 1) / (2 * G))
                if x < ul - cl:
                    out[k] = (rl, ul, pl)
                elif x > um - cml:
                    out[k] = (rl * (pm / pl) ** (1 / G), um, pm)
                else:
                    u = 2 / (G + 1) * (cl + (G - 1) / 2 * ul + x)
                    c = 2 / (G + 1) * (cl + (G - 1) / 2 * (ul - x))
                    rho = rl * (c / cl) ** (2 / (G - 1))
                    out[k] = (rho, u, pl * (c / cl) ** (2 * G / (G - 1)))
        else:
            if pm > pr:
                sr = ur + cr * np.sqrt((G + 1) / (2 * G) * pm / pr + (G - 1) / (2 * G))
                rho = rr * (pm / pr + (G - 1) / (G + 1)) / ((G - 1) / (G + 1) * pm / pr + 1)
                out[k] = (rr, ur, pr) if x > sr else (rho, um, pm)
            else:
                cmr = cr * (pm / pr) ** ((G - 1) / (2 * G))
                if x > ur + cr:
                    out[k] = (rr, ur, pr)
                elif x < um + cmr:
                    out[k] = (rr * (pm / pr) ** (1 / G), um, pm)
                else:
                    u = 2 / (G + 1) * (-cr + (G - 1) / 2 * ur + x)
                    c = 2 / (G + 1) * (cr - (G - 1) / 2 * (ur - x))
                    rho = rr * (c / cr) ** (2 / (G - 1))
                    out[k] = (rho, u, pr * (c / cr) ** (2 * G / (G - 1)))
    return out


@pytest.mark.parametrize("case", ["sod", "toro1"])
def test_euler_converges_to_exact_riemann_solution(case):
    """1D shock tubes along x: the L1 density error against the exact
    solution is small and falls by at least 1.6x when the grid doubles."""
    states = {"sod": (1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 0.2),
              "toro1": (1.0, 0.75, 1.0, 0.125, 0.0, 0.1, 0.2)}[case]
    rl, ul, pl, rr, ur, pr, T = states
    errs = []
    for mx in (100, 200):
        dx = 1.0 / mx
        xc = (np.arange(-1, mx + 3) - 0.5) * dx
        left = xc < 0.3
        q = np.zeros((5, mx + 4, 4))
        q[..., 0] = np.where(left, rl, rr)
        q[..., 1] = np.where(left, rl * ul, rr * ur)
        q[..., 3] = np.where(left, pl / (G - 1) + 0.5 * rl * ul ** 2, pr / (G - 1) + 0.5 * rr * ur ** 2)
        p = euler_problem(mx, 1, mthbc=(1, 1, 1, 1))
        r = O.claw2(p, q, dx, 1.0, 0.0, T, [1e-4, 1e99], [1.0, 0.9], 10000)
        assert r.info == 0 and r.tend == T
        ex = exact_euler_riemann(rl, ul, pl, rr, ur, pr, (xc[2:-2] - 0.3) / T)
        errs.append(np.abs(r.q[2, 2:-2, 0] - ex[:, 0]).mean())
    assert errs[0] < 0.02, errs
    assert errs[0] / errs[1] > 1.6, errs


def test_acoustics_plane_wave_converges():
    """A right-going acoustic pulse p = Z u = f(x - c t) along x on a
    periodic domain: second-order-like convergence of the L1 error."""
    rho, bulk = 1.0, 4.0
    c, Z = 2.0, 2.0
    errs = []
    for mx in (64, 128):
        dx = 1.0 / mx
        xc = (np.arange(-1, mx + 3) - 0.5) * dx
        f = lambda x: np.exp(-80 * (np.mod(x, 1.0) - 0.5) ** 2)  # noqa: E731
        q = np.zeros((5, mx + 4, 3))
        q[..., 0] = f(xc)
        q[..., 1] = f(xc) / Z
        p = acoustics_problem(mx, 1, mthbc=(2, 2, 2, 2), rho=rho, bulk=bulk)
        T = 0.25
        r = O.claw2(p, q, dx, 1.0, 0.0, T, [1e-4, 1e99], [1.0, 0.9], 10000)
        errs.append(np.abs(r.q[2, 2:-2, 0] - f(xc[2:-2] - c * T)).mean())
    assert errs[0] / errs[1] > 3.0, errs
    assert errs[1] < 5e-3, errs


def test_periodic_conservation_oracle():
    q, dx, dy = P.euler_quadrants(40, 30, seed=2)
    p = euler_problem(40, 30, mthbc=(2, 2, 2, 2))
    r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 30)
    t0 = q[2:-2, 2:-2].sum(axis=(0, 1))
    t1 = r.q[2:-2, 2:-2].sum(axis=(0, 1))
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.abs(q[2:-2, 2:-2]).sum(axis=(0, 1)))


def test_info_codes():
    q, dx, dy = P.euler_quadrants(10, 10, seed=1)
    # info = 2: fixed dt does not divide the interval
    p = euler_problem(10, 10, method=(0, 2, 2, 0, 0, 0, 0))
    assert O.claw2(p, q, dx, dy, 0.0, 0.1, [0.03, 1e99], [1.0, 0.9], 100).info == 2
    # info = 3: desired Courant number above the maximum
    p = euler_problem(10, 10)
    assert O.claw2(p, q, dx, dy, 0.0, 0.1, [1e-3, 1e99], [0.5, 0.9], 100).info == 3
    # info = 6: transverse order above method order
    p = euler_problem(10, 10, method=(1, 1, 2, 0, 0, 0, 0))
    assert O.claw2(p, q, dx, dy, 0.0, 0.1, [1e-3, 1e99], [1.0, 0.9], 100).info == 6
    # info = 11: too many steps
    p = euler_problem(10, 10)
    assert O.claw2(p, q, dx, dy, 0.0, 10.0, [1e-3, 1e99], [1.0, 0.9], 3).info == 11


def test_nan_state_stops_the_loop():
    """A non-physical state (seed 0 with a reflecting right wall drives the
    Roe solver to negative pressure) ends the oracle's time loop with
    info = -5 instead of spinning in the retake loop."""
    q, dx, dy = P.euler_quadrants(45, 38, seed=0)
    p = euler_problem(45, 38, mthbc=(1, 3, 2, 2))
    r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 15)
    assert r.info == -5


# --------------------------------------------- aux, capacity: annulus example

def _annulus():
    aux, dxc, dyc = O.annulus_aux(40, 120)
    q = P.annulus_hump(40, 120)
    p = O.Problem(1, 1, 40, 120, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3,
                  method=(1, 2, 2, 0, 0, 3, 3), mthlim=[3], mthbc=(1, 1, 2, 2))
    return p, q, aux, dxc, dyc


def test_golden_annulus_capacity_and_aux():
    """examples/advection_2d_annulus: the oracle's capa/aux path against the
    frames the reference Fortran wrote and the example's regression sums
    (which are round-off: the two humps have opposite signs)."""
    g = np.load(os.path.join(GOLD, "annulus_40x120.npz"))
    p, q, aux, dxc, dyc = _annulus()
    dtv, t = [0.1, 1e9], 0.0
    for k in range(2):
        r = O.claw2(p, q, dxc, dyc, t, t + 0.25, dtv, [1.0, 0.9], 50000, aux=aux)
        assert r.info == 0
        q, dtv, t = r.q, [r.dtv[4], 1e9], r.tend
        assert t == g["t"][k]
        np.testing.assert_array_equal(q[2:-2, 2:-2], g["q"][k])
        assert abs(q[2:-2, 2:-2, 0].sum() - g["regression_sums"][k]) < 1e-10


def test_capacity_weighted_mass_is_conserved_on_the_annulus():
    """With periodic theta and no radial flow (the stream function is
    radial), sum(capa * q) changes only by round-off."""
    p, q, aux, dxc, dyc = _annulus()
    m0 = (q[2:-2, 2:-2, 0] * aux[2:-2, 2:-2, 2]).sum()
    r = O.claw2(p, q, dxc, dyc, 0.0, 0.3, [0.01, 1e9], [1.0, 0.9], 5000, aux=aux)
    m1 = (r.q[2:-2, 2:-2, 0] * aux[2:-2, 2:-2, 2]).sum()
    assert abs(m1 - m0) <= 1e-12 * np.abs(q).sum()


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("method", [(1, 2, 2, 0, 0, 3, 3), (1, 2, 1, 0, 0, 3, 3),
                                    (1, 1, 2, 0, 0, 0, 3), (1, 2, 0, 0, 0, 3, 3)])
def test_live_reference_vc_advection_random(method):
    rng = np.random.default_rng(sum(method))
    mx, my = 23, 17
    aux = np.stack([rng.uniform(-1, 1, (my + 4, mx + 4)), rng.uniform(-1, 1, (my + 4, mx + 4)),
                    rng.uniform(0.5, 1.5, (my + 4, mx + 4))], -1)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3, method=method,
                  mthlim=[4], mthbc=(1, 1, 2, 2))
    q = O.bc2(p, rng.standard_normal((my + 4, mx + 4, 1)))
    a, ca = O.step2(p, q, 0.05, 0.07, 0.01, aux=aux)
    b, cb = O.step2(p, q, 0.05, 0.07, 0.01, aux=aux, backend="ref")
    np.testing.assert_array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2])
    assert ca == cb


@pytest.mark.parametrize("m3", [-1, -2])
def test_golden_annulus_split_with_capacity(m3):
    """The annulus example by dimensional splitting (dimsp2/step2ds with
    aux and capacity) against the reference build's run to t = 0.25."""
    g = np.load(os.path.join(GOLD, "annulus_split_40x120.npz"))
    aux, dxc, dyc = O.annulus_aux(40, 120)
    p = O.Problem(1, 1, 40, 120, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3,
                  method=(1, 2, m3, 0, 0, 3, 3), mthlim=[3], mthbc=(1, 1, 2, 2))
    r = O.claw2(p, P.annulus_hump(40, 120), dxc, dyc, 0.0, 0.25, [0.1, 1e9], [1.0, 0.9], 50000,
                aux=aux)
    key = f"split{-m3}"
    assert r.info == int(g["info_" + key]) == 0 and r.tend == float(g["tend_" + key])
    np.testing.assert_array_equal(r.nv, g["nv_" + key])
    np.testing.assert_array_equal(r.dtv, g["dtv_" + key])
    np.testing.assert_array_equal(r.q[2:-2, 2:-2], g["q_" + key])


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("m3", [-1, -2])
@pytest.mark.parametrize("capa", [0, 3])
@pytest.mark.parametrize("order,lim,fw", [(2, 4, 0), (2, 3, 0), (1, 0, 0), (2, 2, 1)])
def test_live_reference_vc_advection_split(m3, capa, order, lim, fw):
    """One split step (x sweeps over the ghost rows, y over the ghost
    columns, with capacity-scaled dt/dx) against the reference build."""
    rng = np.random.default_rng(10 * -m3 + capa + order + lim)
    mx, my = 23, 17
    aux = np.stack([rng.uniform(-1, 1, (my + 4, mx + 4)), rng.uniform(-1, 1, (my + 4, mx + 4)),
                    rng.uniform(0.5, 1.5, (my + 4, mx + 4))], -1)
    p = O.Problem(1, 1, mx, my, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3,
                  method=(1, order, m3, 0, 0, capa, 3), mthlim=[lim], mthbc=(1, 1, 2, 2),
                  use_fwave=bool(fw))
    q = O.bc2(p, rng.standard_normal((my + 4, mx + 4, 1)))
    a, ca = O.dimsp2(p, q, 0.05, 0.07, 0.01, [1.0, 0.9], aux=aux)
    b, cb = O.dimsp2(p, q, 0.05, 0.07, 0.01, [1.0, 0.9], aux=aux, backend="ref")
    np.testing.assert_array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2])
    assert ca == cb
