"""The numpy oracle of the learned limiter's network: its hand-written
gradients agree with central finite differences, and the clip keeps every
output in the Sweby TVD region with MC outside it (CPU only)."""
import numpy as np

from oracle import mlp_oracle as MO


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(MO.NPARAMS) * 0.3
    X = rng.standard_normal((60, 5))
    th = rng.uniform(-1, 3, 60)
    g, _ = MO.grad_mc_regression(p, X, th)
    dphi = rng.standard_normal(60)
    g2 = MO.grad_from_dphi(p, X, th, dphi)
    eps = 1e-6
    for k in [0, 7, 100, 159, 170, 200, 700, 1215, 1230, 1250, 1280]:
        e = np.zeros_like(p)
        e[k] = eps
        fd = (MO.grad_mc_regression(p + e, X, th)[1] - MO.grad_mc_regression(p - e, X, th)[1]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
        # linear functional sum(dphi * phi) through the clip (inside waves only)
        def lin(pp):
            out, _ = MO.forward(pp, X)
            _, inside = MO.clip(MO.forward(p, X)[0], th)
            return float((dphi * np.where(inside, out, 0.0)).sum())
        fd2 = (lin(p + e) - lin(p - e)) / (2 * eps)
        assert abs(fd2 - g2[k]) <= 1e-6 * max(1.0, abs(fd2)), (k, fd2, g2[k])


def test_clip_region_and_fallback():
    th = np.array([-1.0, 0.0, 0.25, 0.5, 1.0, 2.0, 5.0])
    for raw in (-0.5, 0.0, 0.3, 1.0, 1.9, 2.5):
        phi, inside = MO.clip(np.full_like(th, raw), th)
        hi = np.where(th > 0, np.minimum(2 * th, 2.0), 0.0)
        assert np.all(phi >= 0) and np.all(phi <= hi + 1e-15)
        np.testing.assert_array_equal(phi[~inside], MO.mc(th[~inside]))


def test_classical_targets_known_answers():
    """The regression targets at hand-computed points: superbee is 2 theta up
    to 1/2, 1 up to 1, theta up to 2, then 2 (src/2d/inlinelimiter.f90:78-79);
    MC is 2 theta up to 1/3, (1 + theta) / 2 up to 3, then 2."""
    th = np.array([-1.0, 0.0, 0.2, 0.4, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5, 3.0, 10.0])
    sb = np.array([0.0, 0.0, 0.4, 0.8, 1.0, 1.0, 1.0, 1.5, 2.0, 2.0, 2.0, 2.0])
    mc = np.array([0.0, 0.0, 0.4, 0.7, 0.75, 0.875, 1.0, 1.25, 1.5, 1.75, 2.0, 2.0])
    np.testing.assert_allclose(MO.superbee(th), sb, rtol=0, atol=1e-15)
    np.testing.assert_allclose(MO.mc(th), mc, rtol=0, atol=1e-15)
    # the superbee target's gradient: finite differences of its loss
    rng = np.random.default_rng(1)
    p = rng.standard_normal(MO.NPARAMS) * 0.3
    X = rng.standard_normal((40, 5))
    g, _ = MO.grad_mc_regression(p, X, th.repeat(4)[:40], MO.superbee)
    eps = 1e-6
    for k in [3, 160, 640, 1250, 1280]:
        e = np.zeros_like(p)
        e[k] = eps
        fd = (MO.grad_mc_regression(p + e, X, th.repeat(4)[:40], MO.superbee)[1]
              - MO.grad_mc_regression(p - e, X, th.repeat(4)[:40], MO.superbee)[1]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_shipped_network_loads():
    """The trained parameters shipped with the package: 1281 finite float32
    values whose network (evaluated by the oracle) stays inside the TVD
    region after the clip over a sweep of features."""
    from classic_amd import learned as Lm
    p = Lm.trained_params()
    assert p.shape == (MO.NPARAMS,) and np.isfinite(p).all()
    rng = np.random.default_rng(2)
    X = rng.uniform(-2, 3, (500, 5))
    th = X[:, 0]
    out, _ = MO.forward(p.astype(np.float64), X)
    phi, _ = MO.clip(out, th)
    hi = np.where(th > 0, np.minimum(2 * th, 2.0), 0.0)
    assert np.isfinite(out).all() and np.all(phi >= 0) and np.all(phi <= hi + 1e-15)


def test_evolution_strategies_step_descends():
    """The stage-3 optimiser on a known quadratic: it leaves params[:start]
    alone, its first move points down the gradient, and 200 iterations bring
    the objective close to the minimum it knows nothing about."""
    from classic_amd import learned as Lm
    rng = np.random.default_rng(4)
    target = rng.standard_normal(12)

    def f(p):
        return float(((p[2:] - target[2:]) ** 2).sum())

    p0 = np.zeros(12)
    p0[:2] = [5.0, -3.0]
    best, fs = Lm.es_minimise(f, p0, 2, 200, 6, 0.05, 0.05, np.random.default_rng(0))
    np.testing.assert_array_equal(best[:2], p0[:2])
    assert fs[1] < fs[0]
    assert min(fs) <= 1e-2 * fs[0] and f(best) == min(fs)


# ------------------------------------------------- the feature restatement

def _euler_state(mx, my, seed, mthbc=(1, 1, 1, 1)):
    from classic_amd import problems as P
    from oracle import oracle as O

    q, dx, dy = P.euler_quadrants(mx, my, seed=seed)
    p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                  mthlim=[4, 4, 4], mthbc=mthbc)
    r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 6)  # evolved, waves everywhere
    return O.bc2(p, r.q), dx, dy


def test_features_theta_is_the_reference_limiter_ratio():
    """theta of every nonzero wave is the r of the reference's limiter: the
    oracle limiter (bit for bit with src/2d/inlinelimiter.f90) with the
    Beam-Warming choice phi = r turns each wave W into exactly r W."""
    from oracle import oracle as O

    mx, my = 23, 17
    q, dx, dy = _euler_state(mx, my, seed=4)
    par = O.solver_params(gamma=1.4, efix=True)
    feat, theta, nz = MO.features(q, dx, dy, 0.3 * dx, par)
    checked = 0
    for d in (0, 1):
        n, nslice = (mx, my + 2) if d == 0 else (my, mx + 2)
        for p in range(nslice):
            cells = q[p + 1] if d == 0 else q[:, p + 1]
            wave, s, _, _ = O.rpn2(O.SOLVER_EULER, par, d + 1, cells, 3)
            lim = O.limiter(wave, s, [5, 5, 5])
            for k in range(1, n + 2):
                e = k + 1
                slot = (p + 1) * (mx + 4) + (k + 1) if d == 0 else (k + 1) * (mx + 4) + (p + 1)
                for mw in range(3):
                    if nz[d, slot, mw]:
                        np.testing.assert_array_equal(lim[e, mw], theta[d, slot, mw] * wave[e, mw])
                        checked += 1
                    else:
                        assert not np.any(wave[e, mw])
    assert checked > 1000


def test_features_of_a_uniform_state_are_zero():
    from oracle import oracle as O

    q = np.zeros((10, 12, 4))
    q[..., 0], q[..., 1], q[..., 2], q[..., 3] = 1.3, 0.2, -0.1, 2.5
    feat, theta, nz = MO.features(q, 0.1, 0.1, 0.01, O.solver_params(gamma=1.4, efix=True))
    assert not nz.any() and not feat.any() and not theta.any()


def _advected_density(mx, my, u, rho_of_i):
    """Density varying along x, constant velocity (u, 0) and pressure 1: the
    Roe solver sees pure entropy waves W = (drho, u drho, 0, u^2 drho / 2)."""
    q = np.zeros((my + 4, mx + 4, 4))
    rho = np.array([rho_of_i(i) for i in range(-1, mx + 3)])
    q[..., 0] = rho[None, :]
    q[..., 1] = rho[None, :] * u
    q[..., 3] = 1.0 / 0.4 + 0.5 * rho[None, :] * u * u
    return q


def test_features_known_answers_quadratic_entropy_profile():
    """rho_i = 2 + i^2/64: jumps drho_k = (2k - 1)/64 across interface k, so
    with u > 0 the upwind ratio is (2k-3)/(2k-1) and the downwind one
    (2k+1)/(2k-1); with u < 0 they swap.  The middle wave carries it all:
    f2 = u dt/dx, f3 = |W|/(|q_l| + |q_r|), f4 = u / max|u -+ a|, and the
    acoustic waves vanish to rounding."""
    from oracle import oracle as O

    mx, my, dx, dt = 12, 3, 0.05, 0.01
    par = O.solver_params(gamma=1.4, efix=True)
    for u in (0.3, -0.3):
        q = _advected_density(mx, my, u, lambda i: 2.0 + i * i / 64.0)
        feat, theta, nz = MO.features(q, dx, dx, dt, par)
        for J in range(0, my + 2):
            for k in range(2, mx + 1):
                slot = (J + 1) * (mx + 4) + (k + 1)
                # the acoustic waves vanish up to the rounding of the Roe averages
                assert nz[0, slot, 1]
                assert feat[0, slot, 0, 3] <= 1e-12 and feat[0, slot, 2, 3] <= 1e-12
                up, down = (2 * k - 3) / (2 * k - 1), (2 * k + 1) / (2 * k - 1)
                if u < 0:
                    up, down = down, up
                assert abs(theta[0, slot, 1] - up) <= 1e-13
                f = feat[0, slot, 1].astype(np.float64)
                assert abs(f[0] - min(up, 3.0) / 3.0) <= 1e-7
                assert abs(f[1] - min(down, 3.0) / 3.0) <= 1e-7
                assert abs(f[2] - u * dt / dx) <= 1e-7 * abs(u * dt / dx)
                ql, qr = q[J + 1, k], q[J + 1, k + 1]  # cells k-1 and k
                w = qr - ql
                want = np.linalg.norm(w) / (np.linalg.norm(ql) + np.linalg.norm(qr))
                assert abs(f[3] - want) <= 1e-6 * want
                # u / max(|u - a|, |u|, |u + a|) = u / (|u| + a), a the Roe
                # sound speed of the interface (FVMHP 15.3.2)
                sl, sr = np.sqrt(ql[0]), np.sqrt(qr[0])
                hl = (ql[3] + 1.0) / ql[0]
                hr = (qr[3] + 1.0) / qr[0]
                h = (sl * hl + sr * hr) / (sl + sr)
                a = np.sqrt(0.4 * (h - 0.5 * u * u))
                assert abs(f[4] - u / (abs(u) + a)) <= 1e-6
        # the y-interfaces see no jump
        assert not nz[1].any()


def test_features_known_answers_isolated_jump():
    """A single density jump (u = v = 0, constant pressure): one nonzero
    wave, the entropy wave at the jump, with both ratios 0 (its neighbours
    are zero waves), f2 = f4 = 0 (zero speed) and f3 = |drho| / (|q_l| +
    |q_r|)."""
    from oracle import oracle as O

    mx, my = 8, 2
    q = _advected_density(mx, my, 0.0, lambda i: 1.0 if i <= 4 else 0.25)
    feat, theta, nz = MO.features(q, 0.1, 0.1, 0.01, O.solver_params(gamma=1.4, efix=True))
    where = np.argwhere(nz[0])
    assert len(where) == my + 2  # one wave per row, at interface 5
    for slot, mw in where:
        assert mw == 1 and slot % (mx + 4) == 6
        f = feat[0, slot, 1].astype(np.float64)
        ql = np.array([1.0, 0.0, 0.0, 2.5])
        qr = np.array([0.25, 0.0, 0.0, 2.5])
        want = 0.75 / (np.linalg.norm(ql) + np.linalg.norm(qr))
        assert f[0] == 0.0 and f[1] == 0.0 and f[2] == 0.0 and f[4] == 0.0
        assert abs(f[3] - want) <= 1e-6 * want
        assert theta[0, slot, 1] == 0.0
