"""Learned limiter (limiter 6): the table path, the BF16 MFMA inference, the
FP32 training gradients against the numpy oracle, and what training gains
over the plain MC limiter on held-out data."""
import numpy as np
import pytest

from classic_amd import SOLVER_EULER
from classic_amd import learned as Lm
from classic_amd import problems as P
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu

METHOD, BC = (1, 2, 2, 0, 0, 0, 0), (1, 1, 1, 1)


def _fallback_params():
    """Network output -1 everywhere: outside the TVD region, so MC is used."""
    p = np.zeros(Lm.NPARAMS, np.float32)
    p[-1] = -1.0
    return p


def _state(n=96, seed=3, steps=10):
    q, dx, dy = P.euler_quadrants(n, n, seed=seed)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0]) as g:
        g.set_q(q)
        g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], steps, METHOD, [4, 4, 4], BC)
        return g.get_q(), dx, dy


@pytest.mark.parametrize("arith", [0, 1])
def test_table_limiter_with_mc_fallback_equals_mc(arith):
    """With every network output outside the TVD region the table holds the
    MC values: the step is identical to limiter 4 -- bit for bit in exact
    arithmetic, where the pre-pass computes theta like the step kernel."""
    q, dx, dy = _state()
    n = q.shape[0] - 4
    res = []
    for lim in (4, 6):
        with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith) as g:
            g.set_params(_fallback_params())
            g.set_q(q)
            out = g.claw2(dx, dy, 0.0, 1e9, [0.2 * dx, 1e99], [1.0, 0.9], 5, METHOD, [lim] * 3, BC)
            res.append((g.get_q(), out["dtv"].copy()))
    if arith == 0:
        np.testing.assert_array_equal(res[0][0][2:-2, 2:-2], res[1][0][2:-2, 2:-2])
        np.testing.assert_array_equal(res[0][1], res[1][1])
    else:  # the pre-pass rounds theta without FMA, the fast kernel with it
        from tests.helpers import max_rel_err
        assert max_rel_err(res[1][0], res[0][0]) <= 1e-13


def _oracle_features(q, dx, dy, dt):
    from oracle import oracle as O

    return MO.features(q, dx, dy, dt, O.solver_params(gamma=1.4, efix=True))


def _device_features(q, dx, dy, dt, arith):
    n = q.shape[0] - 4
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith) as g:
        g.set_params(Lm.init_params(7))
        g.set_q(q)
        g.bc2(BC)
        qb = g.get_q()  # ghost cells filled by the device bc2 (bit for bit the oracle's)
        g.features(dx, dy, dt)
        nw = g.nwaves()
        ntab = nw // 3
        feat = g.read(0).reshape(2, ntab, 3, 8)[..., :5]
        theta = g.read(3).reshape(2, ntab, 3)
    return qb, feat, theta


@pytest.mark.parametrize("arith", [0, 1])
def test_features_vs_oracle(arith):
    """The feature pre-pass (slot_features: three Riemann solves per table
    slot) against the oracle's restatement on the 96^2 evolved state.  Exact
    arithmetic: theta bit for bit, the float features equal.  Fast
    arithmetic: see _fast_close."""
    q, dx, dy = _state()
    dt = 0.2 * dx
    qb, feat, theta = _device_features(q, dx, dy, dt, arith)
    ofeat, oth, onz = _oracle_features(qb, dx, dy, dt)
    assert onz.sum() > 5000
    live = feat[..., 3] > 0
    np.testing.assert_array_equal(live, onz & (ofeat[..., 3] > 0))
    if arith == 0:
        np.testing.assert_array_equal(theta, oth)
        np.testing.assert_array_equal(feat, ofeat)
    else:
        _fast_close(theta, oth, feat, ofeat)


def _fast_close(theta, oth, feat, ofeat):
    """Fast arithmetic computes a wave of strength eps from O(1) states to
    about 1e-16 absolute, so theta, a ratio of the wave's dot products, is
    only as good as 1e-16 / eps relatively; and beyond +-3 nothing uses it
    (the feature is clamped there, MC is 2 and the TVD bound is 2).  So theta
    is compared clamped to [-3, 3] on waves whose strength relative to the
    two states (feature 3) is at least 1e-6: within 1e-9; the float features
    of every wave within 1e-6."""
    sig = ofeat[..., 3] >= 1e-6
    ct, co = np.clip(theta, -3.0, 3.0), np.clip(oth, -3.0, 3.0)
    dth = np.abs(ct - co)[sig] / np.maximum(np.abs(co[sig]), 1.0)
    d = np.abs(feat.astype(np.float64) - ofeat.astype(np.float64)) / np.maximum(np.abs(ofeat), 1.0)
    print(f"theta ({int(sig.sum())} significant waves): max {dth.max():.2e}; "
          f"features: max {d.max():.2e}", flush=True)
    assert dth.max() <= 1e-9
    assert d.max() <= 1e-6


def test_features_vs_oracle_1024():
    """The same at the learned workload's size (1024^2, a few steps into the
    four-quadrant run), fast arithmetic as the benchmark runs it."""
    n = 1024
    q, dx, dy = P.euler_quadrants(n, n, seed=200)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=1) as g:
        g.set_q(q)
        g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 8, METHOD, [4, 4, 4], BC)
        q = g.get_q()
    dt = 0.25 * dx
    qb, feat, theta = _device_features(q, dx, dy, dt, 1)
    ofeat, oth, onz = _oracle_features(qb, dx, dy, dt)
    assert onz.sum() > 50_000
    _fast_close(theta, oth, feat, ofeat)


@pytest.mark.parametrize("arith", [0, 1])
def test_fused_prepass_phi_vs_oracle_network(arith):
    """The phi tables a limiter-6 step computes (fused pre-pass: features,
    BF16 MFMA network, TVD clip with MC fallback) against the oracle network
    on the oracle's features: within BF16 rounding, except where the clip
    decision flips at the region's edge; zero waves get phi = 0."""
    q, dx, dy = _state()
    n = q.shape[0] - 4
    dt = 0.2 * dx
    params = Lm.init_params(7)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith) as g:
        g.set_params(params)
        g.set_q(q)
        g.bc2(BC)
        qb = g.get_q()
        g.step2(dx, dy, dt, METHOD, [6, 6, 6])
        phi = np.concatenate([g.read(1), g.read(2)]).reshape(2, -1, 3)
    ofeat, oth, onz = _oracle_features(qb, dx, dy, dt)
    live = onz & (ofeat[..., 3] > 0)
    raw, _ = MO.forward(params.astype(np.float64), ofeat[live].astype(np.float64))
    ref, _ = MO.clip(raw, oth[live])
    d = np.abs(phi[live] - ref)
    assert np.median(d) < 1e-2 and np.mean(d < 5e-2) > 0.99, (np.median(d), np.mean(d < 5e-2))
    assert not phi[~onz].any()


def _tables(g):
    nw = g.nwaves()
    return g.read(0), g.read(3), np.concatenate([g.read(1), g.read(2)]), nw


def test_inference_fp32_and_mfma_vs_oracle():
    q, dx, dy = _state()
    n = q.shape[0] - 4
    params = Lm.init_params(7)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0]) as g:
        g.set_params(params)
        g.set_q(q)
        g.bc2(BC)
        g.features(dx, dy, 0.2 * dx)
        g.phi(1)
        feat, theta, phi32, nw = _tables(g)
        g.phi(0)
        phibf = np.concatenate([g.read(1), g.read(2)])
    live = feat[:, 3] > 0
    assert live.sum() > 1000
    raw, _ = MO.forward(params.astype(np.float64), feat[live, :5].astype(np.float64))
    ref, inside = MO.clip(raw, theta[live])
    # FP32 network vs float64 oracle
    d32 = np.abs(phi32[live] - ref)
    assert np.quantile(d32, 0.999) < 1e-5, d32.max()
    # BF16 matrix-core network: within BF16 rounding except where the clip
    # decision flips at the region boundary
    dbf = np.abs(phibf[live] - ref)
    assert np.median(dbf) < 1e-2 and np.mean(dbf < 5e-2) > 0.99, (np.median(dbf), np.mean(dbf < 5e-2))
    # waves without a nonzero wave keep the MC default
    np.testing.assert_array_equal(phibf[~live], phi32[~live])


@pytest.mark.parametrize("arith", [0, 1])
def test_fused_prepass_equals_features_then_inference(arith):
    """A step with limiter 6 computes its phi tables in one fused kernel
    (features through LDS into the MFMA network): the tables equal those of
    the separate feature pass and MFMA inference -- bit for bit in exact
    arithmetic; in fast arithmetic the compiler may contract the features'
    FP64 expressions differently in the two kernels, which moves a float
    feature by an ulp now and then: tables within float rounding, a BF16
    rounding step at most."""
    q, dx, dy = _state()
    n = q.shape[0] - 4
    dt = 0.2 * dx
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith) as g:
        g.set_params(Lm.init_params(7))
        # both passes must write every entry: start from tables of NaN
        nw = g.nwaves()
        g.write(1, np.full(nw, np.nan))
        g.write(2, np.full(nw, np.nan))
        g.set_q(q)
        g.bc2(BC)
        g.step2(dx, dy, dt, METHOD, [6, 6, 6])
        fused = np.concatenate([g.read(1), g.read(2)])
        # fixed-dt stepping runs the same pre-pass
        g.write(1, np.full(nw, np.nan))
        g.write(2, np.full(nw, np.nan))
        g.set_q(q)
        g.time_steps(dx, dy, dt, METHOD, [6, 6, 6], BC, 1)
        timed = np.concatenate([g.read(1), g.read(2)])
        g.write(1, np.full(nw, np.nan))
        g.write(2, np.full(nw, np.nan))
        g.set_q(q)
        g.bc2(BC)
        g.features(dx, dy, dt)
        g.phi(0)
        feat = g.read(0)
        twopass = np.concatenate([g.read(1), g.read(2)])
    live = feat[:, 3] > 0
    assert live.sum() > 1000 and np.any(fused[live] != 0.0)
    assert not np.isnan(fused).any() and not np.isnan(twopass).any()
    np.testing.assert_array_equal(timed, fused)
    if arith == 0:
        np.testing.assert_array_equal(fused, twopass)
    else:
        d = np.abs(fused - twopass)
        close = d <= 1e-6 * np.abs(twopass) + 1e-300
        assert np.mean(close) > 0.999 and d.max() < 5e-2, (np.mean(close), d.max())


@pytest.mark.parametrize("solver", ["euler", "acoustics"])
@pytest.mark.parametrize("arith", [0, 1])
def test_prepass_ragged_grids_equal_features_then_inference(solver, arith):
    """The pre-pass takes lines of 61 table slots per wavefront (table rows for
    x, table columns for y), one Riemann solve per interface, and queues only
    the waves that need the network: on grids whose sides are not multiples
    of 61 (130 x 45 Euler, 75 x 123 acoustics, both evolved) its tables equal
    the separate feature pass (three solves per slot, every wave) followed by
    the MFMA inference over all waves -- bit for bit, every entry written."""
    from classic_amd import SOLVER_ACOUSTICS

    if solver == "euler":
        mx, my, sid, prm, lim = 130, 45, SOLVER_EULER, [1.4, 1.0], [4, 4, 4]
        q, dx, dy = P.euler_quadrants(mx, my, seed=5)
        bc = BC
    else:
        mx, my, sid, prm, lim = 75, 123, SOLVER_ACOUSTICS, [1.0, 4.0], [4, 4]
        q, dx, dy = P.acoustics_pulse(mx, my)
        bc = (3, 3, 3, 3)
    with Lm.LearnedGrid(sid, mx, my, prm, arith=arith) as g:
        g.set_q(q)
        g.claw2(dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 12, METHOD, lim, bc)
        q = g.get_q()
    dt = 0.2 * dx
    with Lm.LearnedGrid(sid, mx, my, prm, arith=arith) as g:
        g.set_params(Lm.init_params(11))
        nw = g.nwaves()
        g.write(1, np.full(nw, np.nan))
        g.write(2, np.full(nw, np.nan))
        g.set_q(q)
        g.bc2(bc)
        g.step2(dx, dy, dt, METHOD, [6] * g.mwaves)
        march = np.concatenate([g.read(1), g.read(2)])
        g.write(1, np.full(nw, np.nan))
        g.write(2, np.full(nw, np.nan))
        g.set_q(q)  # the step advanced the device q
        g.bc2(bc)
        g.features(dx, dy, dt)
        g.phi(0)
        feat = g.read(0)
        theta = g.read(3)
        twopass = np.concatenate([g.read(1), g.read(2)])
    need = (feat[:, 3] > 0) & (theta > 0)
    print(f"{solver}: {need.sum()} of {need.size} waves through the network", flush=True)
    assert need.sum() > 100 and not np.isnan(march).any()
    np.testing.assert_array_equal(march, twopass)


def test_timed_passes_leave_the_inference_tables():
    """The measurement entry point runs the fused pre-pass, a feature pass
    and the MFMA inference; its phi tables equal one feature + inference
    pass."""
    q, dx, dy = _state()
    n = q.shape[0] - 4
    params = Lm.init_params(7)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0]) as g:
        g.set_params(params)
        g.set_q(q)
        g.bc2(BC)
        g.features(dx, dy, 0.2 * dx)
        g.phi(0)
        want = np.concatenate([g.read(1), g.read(2)])
        g.write(1, np.full(g.nwaves(), np.nan))
        g.write(2, np.full(g.nwaves(), np.nan))
        f_ms, i_ms = g.time_kernels(dx, dy, 0.2 * dx, 3)
        got = np.concatenate([g.read(1), g.read(2)])
    assert f_ms > 0.0 and i_ms > 0.0
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode,target", [(0, MO.mc), (2, MO.superbee)])
def test_mc_regression_gradient_vs_oracle(mode, target):
    """Modes 0 and 2: regression onto MC and onto superbee."""
    q, dx, dy = _state(n=64)
    n = 64
    params = Lm.init_params(3)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0]) as g:
        g.set_params(params)
        g.set_q(q)
        g.bc2(BC)
        g.features(dx, dy, 0.2 * dx)
        gr, loss = g.grad(dx, dy, 0.2 * dx, mode)
        feat, theta = g.read(0), g.read(3)
    live = feat[:, 3] > 0
    go, lo = MO.grad_mc_regression(params.astype(np.float64), feat[live, :5].astype(np.float64),
                                   theta[live], target)
    assert abs(loss - lo) <= 1e-4 * lo
    np.testing.assert_allclose(gr, go, rtol=2e-3, atol=1e-4 * np.abs(go).max())


@pytest.mark.parametrize("mode", [1, 3])
def test_one_step_gradient_vs_oracle(mode):
    """Modes 1 and 3 (squared and absolute one-step error): the dL/dphi table
    written by the kernel, back-propagated by the numpy oracle through the
    same clip, gives the kernel's parameter gradient."""
    n = 64
    snaps, dxc, dtc = Lm.fine_snapshots(n, seed=21, nsnap=1)
    a, b = snaps[0]
    params = Lm.init_params(5)
    with Lm.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=1) as g:
        g.set_params(params)
        g.set_q(a)
        g.set_target(b)
        gr, loss = g.grad(dxc, dxc, dtc, mode)
        feat, theta, dphi = g.read(0), g.read(3), g.read(5)
    live = feat[:, 3] > 0
    go = MO.grad_from_dphi(params.astype(np.float64), feat[live, :5].astype(np.float64),
                           theta[live], dphi[live].astype(np.float64))
    assert loss > 0
    np.testing.assert_allclose(gr, go, rtol=5e-3, atol=1e-3 * np.abs(go).max())


def test_training_gains_over_mc_on_held_out_data():
    """Train briefly on 4x-coarsened 512^2 runs (3 seeds) and compare the
    one-step error of learned vs MC on held-out seeds; the trained network
    must not be worse than MC and the multi-step run must stay finite."""
    params, hist = Lm.train(n_coarse=128, seeds=(31, 32, 33), nsnap=3, pre_iters=150, iters=200,
                            lr=2e-3)
    assert np.isfinite(hist["step"]).all()
    held = []
    with Lm.LearnedGrid(SOLVER_EULER, 128, 128, [1.4, 1.0], arith=1) as g:
        for s in (41, 42):
            snaps, dx, dtc = Lm.fine_snapshots(128, seed=s, nsnap=3)
            for a, b in snaps:
                mc = Lm.one_step_loss(g, a, b, dx, dtc, None)
                ln = Lm.one_step_loss(g, a, b, dx, dtc, params)
                held.append((mc, ln))
    held = np.array(held)
    gain = held[:, 0].sum() / held[:, 1].sum()
    print("held-out one-step error MC / learned =", gain, "pre-train loss", hist["pre"][-1],
          "train loss first/last", hist["step"][0], hist["step"][-1])
    assert gain >= 1.0
    q, el, steps = Lm.run(128, 41, 0.05, 6, params)
    assert np.isfinite(q).all() and steps > 0


def test_shipped_network_beats_mc_over_many_steps():
    """The shipped network on a held-out seed at 1024^2 to t = 0.1 (about 310
    steps at CFL 0.9): its L1 error against the 4x-fine plain run is below
    plain MC's at the same resolution, and the run is finite and takes as
    many steps as MC within 5 %."""
    n, seed, tend = 1024, 200, 0.1
    qf, _, _ = Lm.run(4 * n, seed, tend, 4)
    ref = Lm.coarsen(qf, 4)
    del qf
    ql, _, nl = Lm.run(n, seed, tend, 6, Lm.trained_params())
    qm, _, nm = Lm.run(n, seed, tend, 4)
    l1l, l1m = Lm.l1_error(ql, ref), Lm.l1_error(qm, ref)
    print(f"seed {seed}: L1 learned {l1l:.3e} ({nl} steps), MC {l1m:.3e} ({nm} steps)")
    assert np.isfinite(ql).all()
    assert abs(nl - nm) <= 0.05 * nm
    assert l1l < 0.75 * l1m
