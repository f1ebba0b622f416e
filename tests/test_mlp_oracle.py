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
