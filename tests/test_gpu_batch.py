"""Batched 2D path (clawgpu_batch_*) against single-grid claw2 runs.

Every problem of a batch follows its own claw2 loop (src/2d/claw2.f:494-784)
on the device; the batched result of each problem must equal, bit for bit,
what clawgpu_claw2 produces for that grid alone (and, for small grids, what
the CPU oracle's claw2 produces), including the dt sequence, retakes,
Courant numbers and the info codes.  The schedule (tile height, grid size,
tile order), graph replay and autotuning must not change any bit.  In FAST
arithmetic the single-grid loop kernel (ghost cells read through the
boundary maps) and the batch kernel are different instantiations whose
multiply-adds the compiler may fuse differently, so there the bar is the
FAST bar: 1e-12 relative, the same step counts and info codes.
"""
import ctypes as C

import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from classic_amd import problems as P
from classic_amd.batch import Batch2D, from_specs
from oracle import oracle as O
from tests.helpers import interior, max_rel_err, smooth_acoustics

pytestmark = pytest.mark.gpu

METHOD = (1, 2, 2, 0, 0, 0, 0)
MTHLIM = (4, 4, 4)

# ragged sizes, both solvers, every boundary kind
SIZES = [(33, 47, "euler"), (61, 60, "acoustics"), (120, 37, "euler"), (17, 90, "acoustics"),
         (64, 64, "euler"), (130, 21, "acoustics"), (7, 5, "euler"), (3, 11, "acoustics"),
         (1, 1, "euler")]
BCS = [(1, 1, 1, 1), (3, 3, 3, 3), (2, 2, 1, 3), (3, 1, 2, 2), (1, 3, 3, 1), (2, 2, 2, 2),
       (1, 1, 1, 1), (3, 3, 3, 3), (1, 1, 1, 1)]


def _specs(seed=0, sizes=SIZES, bcs=BCS):
    rng = np.random.default_rng(seed)
    specs, qs = [], []
    for k, ((mx, my, kind), bc) in enumerate(zip(sizes, bcs)):
        if kind == "euler":
            q, dx, dy = P.euler_quadrants(mx, my, seed=seed * 100 + k)
            params = (1.4, 1.0)
        else:
            q = smooth_acoustics(mx, my, seed=seed * 100 + k)
            dx, dy = 1.0 / mx, 1.0 / my
            params = (rng.uniform(0.5, 4.0), rng.uniform(1.0, 8.0))
        specs.append(dict(solver=kind, mx=mx, my=my, params=params, mthbc=bc, dx=dx, dy=dy))
        qs.append(q)
    return specs, qs


def _single(spec, q, tend, dtv, cflv, nmax, method=METHOD, mthlim=MTHLIM, arith=0):
    sid = SOLVER_EULER if spec["solver"] == "euler" else SOLVER_ACOUSTICS
    mw = 3 if sid == SOLVER_EULER else 2
    with gpu.Grid2D(sid, spec["mx"], spec["my"], list(spec["params"]), arith=arith) as g:
        g.set_q(q)
        out = g.claw2(spec["dx"], spec["dy"], 0.0, tend, dtv, cflv, nmax, method, mthlim[:mw],
                      spec["mthbc"])
        return g.get_q(), out


def _batch(specs, qs, tend, dtv, cflv, nmax, method=METHOD, mthlim=MTHLIM, arith=0, graph=True,
           schedule=None, tune=False):
    with Batch2D(specs, arith=arith) as b:
        for k, q in enumerate(qs):
            b.set_q(k, q)
        if schedule:
            b.set_schedule(*schedule)
        b.start(0.0, tend, dtv, cflv, nmax, method, mthlim)
        if tune:
            b.advance(3, graph=graph)
            cands = b.tune(reps=1)
            assert len(cands) == 112  # 7 tile heights x 4 grid sizes x 2 orders x 2 budgets
            best = min(cands, key=lambda c: c[4])
            s = b.schedule()
            assert (s["seglen"], s["nblocks"], s["largest_first"], s["w3"]) == best[:4]
        left = b.advance(graph=graph)
        assert left == 0
        return [b.get_q(k) for k in range(len(specs))], [b.result(k) for k in range(len(specs))]


def _same_outputs(r, s):
    assert r["info"] == s["info"]
    assert r["tend"] == s["tend"]
    assert r["nv"][1] == s["nv"][1]
    np.testing.assert_array_equal(r["dtv"][2:5], s["dtv"][2:5])
    np.testing.assert_array_equal(r["cflv"][2:4], s["cflv"][2:4])


def _tends(n, lo=0.02, hi=0.12, seed=1):
    return np.random.default_rng(seed).uniform(lo, hi, n)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_grid_claw2(graph):
    specs, qs = _specs(0)
    tend = _tends(len(specs))
    dtv, cflv = [1e-3, 1e99], [1.0, 0.9]
    qb, rb = _batch(specs, qs, tend, dtv, cflv, 5000, graph=graph)
    for k, spec in enumerate(specs):
        qg, rg = _single(spec, qs[k], tend[k], dtv, cflv, 5000)
        print("problem", k, spec["solver"], spec["mx"], spec["my"], "steps", rb[k]["nv"][1],
              flush=True)
        np.testing.assert_array_equal(interior(qb[k]), interior(qg))
        _same_outputs(rb[k], rg)
        assert rb[k]["info"] == 0 and rb[k]["tend"] >= tend[k]


def test_batch_retakes_and_maxn():
    """dt starts far too large (every problem retakes, claw2.f:695-720) and
    nv(1) caps the step count (info = 11 where tend is not reached)."""
    specs, qs = _specs(1)
    tend = _tends(len(specs), 0.05, 0.5, seed=2)
    dtv, cflv = [0.2, 1e99], [1.0, 0.9]
    qb, rb = _batch(specs, qs, tend, dtv, cflv, 12)
    infos = set()
    for k, spec in enumerate(specs):
        qg, rg = _single(spec, qs[k], tend[k], dtv, cflv, 12)
        np.testing.assert_array_equal(interior(qb[k]), interior(qg))
        _same_outputs(rb[k], rg)
        infos.add(rb[k]["info"])
    assert 11 in infos


def test_batch_fixed_dt_info_codes():
    """method(1) = 0: fixed dt; a step over cflv(1) ends that problem with
    info = 12 (claw2.f:771-775), a tend not a multiple of dt gives info = 2."""
    specs, qs = _specs(2)
    n = len(specs)
    dtv, cflv = [4e-3, 1e99], [1.0, 0.9]
    tend = np.full(n, 0.04)
    tend[3] = 0.0415  # not a multiple of dt
    method = (0, 2, 2, 0, 0, 0, 0)
    qb, rb = _batch(specs, qs, tend, dtv, cflv, 100, method=method)
    infos = []
    for k, spec in enumerate(specs):
        qg, rg = _single(spec, qs[k], tend[k], dtv, cflv, 100, method=method)
        np.testing.assert_array_equal(interior(qb[k]), interior(qg))
        _same_outputs(rb[k], rg)
        infos.append(rb[k]["info"])
    assert infos[3] == 2 and 12 in infos and 0 in infos


@pytest.mark.parametrize("schedule", [(16, 0, True, False), (64, 256, False, True),
                                      (128, 1024, True, True), (5, 64, False, False)])
def test_batch_schedule_does_not_change_results(schedule):
    specs, qs = _specs(3)
    tend = _tends(len(specs), seed=3)
    dtv, cflv = [1e-3, 1e99], [1.0, 0.9]
    q0, r0 = _batch(specs, qs, tend, dtv, cflv, 5000)
    q1, r1 = _batch(specs, qs, tend, dtv, cflv, 5000, schedule=schedule)
    for k in range(len(specs)):
        np.testing.assert_array_equal(interior(q0[k]), interior(q1[k]))
        _same_outputs(r0[k], r1[k])


def test_batch_tune_preserves_state():
    """Autotuning runs dry iterations mid-run; the outcome is unchanged."""
    specs, qs = _specs(4)
    tend = _tends(len(specs), seed=4)
    dtv, cflv = [1e-3, 1e99], [1.0, 0.9]
    q0, r0 = _batch(specs, qs, tend, dtv, cflv, 5000)
    q1, r1 = _batch(specs, qs, tend, dtv, cflv, 5000, tune=True)
    for k in range(len(specs)):
        np.testing.assert_array_equal(interior(q0[k]), interior(q1[k]))
        _same_outputs(r0[k], r1[k])


def test_batch_fast_arith_equals_single_fast():
    """FAST arithmetic runs the single-grid FAST wave code per tile: the same
    results to the FAST bar, the same steps and info codes."""
    specs, qs = _specs(5)
    tend = _tends(len(specs), seed=5)
    dtv, cflv = [1e-3, 1e99], [1.0, 0.9]
    qb, rb = _batch(specs, qs, tend, dtv, cflv, 5000, arith=1)
    for k, spec in enumerate(specs):
        qg, rg = _single(spec, qs[k], tend[k], dtv, cflv, 5000, arith=1)
        assert max_rel_err(qb[k], qg) <= 1e-12
        assert rb[k]["info"] == rg["info"] and rb[k]["nv"][1] == rg["nv"][1]
        assert rb[k]["tend"] == rg["tend"]  # both end on tend exactly
        np.testing.assert_allclose(rb[k]["dtv"][2:5], rg["dtv"][2:5], rtol=1e-12)
        np.testing.assert_allclose(rb[k]["cflv"][2:4], rg["cflv"][2:4], rtol=1e-12)


def test_batch_matches_oracle_claw2():
    """Two small problems straight against the CPU oracle's claw2."""
    specs, qs = _specs(6, sizes=[(40, 31, "euler"), (29, 44, "acoustics")],
                       bcs=[(1, 1, 1, 1), (3, 3, 3, 3)])
    tend = [0.06, 0.08]
    dtv, cflv = [1e-3, 1e99], [1.0, 0.9]
    qb, rb = _batch(specs, qs, tend, dtv, cflv, 5000)
    for k, spec in enumerate(specs):
        if spec["solver"] == "euler":
            p = O.Problem(4, 3, spec["mx"], spec["my"], O.SOLVER_EULER,
                          O.solver_params(gamma=1.4, efix=True), method=METHOD,
                          mthlim=[4, 4, 4], mthbc=spec["mthbc"])
        else:
            p = O.Problem(3, 2, spec["mx"], spec["my"], O.SOLVER_ACOUSTICS,
                          O.solver_params(*spec["params"]), method=METHOD, mthlim=[4, 4],
                          mthbc=spec["mthbc"])
        r = O.claw2(p, qs[k], spec["dx"], spec["dy"], 0.0, tend[k], dtv, cflv, 5000)
        np.testing.assert_array_equal(interior(qb[k]), interior(r.q))
        assert rb[k]["nv"][1] == r.nv[1] and rb[k]["tend"] == r.tend


def test_batch_256_problem_workload_sample():
    """The 256-problem workload (sides log-uniform in [128, 1024]): after 6
    batched iterations in FAST arithmetic, the smallest, a middle and the
    largest problem agree with their single-grid runs of the same steps to
    the FAST bar."""
    specs = P.batch_problem_specs(256, seed=0)
    b, qs = from_specs(specs, arith=1)
    with b:
        b.start(0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 10**6, METHOD, MTHLIM)
        b.tune(reps=1)
        b.advance(6)
        cells = [s["mx"] * s["my"] for s in specs]
        order = np.argsort(cells)
        for k in (int(order[0]), int(order[128]), int(order[-1])):
            r = b.result(k)
            qb = b.get_q(k)
            spec = dict(specs[k])
            _, dx, dy = P.batch_problem_q(spec)
            spec.update(dx=dx, dy=dy)
            qg, rg = _single(spec, qs[k], 1e9, [1e-4, 1e99], [1.0, 0.9], int(r["nv"][1]), arith=1)
            print("problem", k, spec["solver"], spec["mx"], spec["my"], "steps", r["nv"][1],
                  flush=True)
            assert r["nv"][1] >= 1
            assert max_rel_err(qb, qg) <= 1e-12
            assert abs(r["tend"] - rg["tend"]) <= 1e-12 * r["tend"]


def test_batch_of_one_and_zero_steps():
    """One problem equals its single-grid run; nv(1) = 0 finishes every
    problem at once with the initial data untouched."""
    specs, qs = _specs(7, sizes=[(45, 38, "euler")], bcs=[(1, 1, 1, 1)])
    qb, rb = _batch(specs, qs, [0.05], [1e-3, 1e99], [1.0, 0.9], 5000)
    qg, rg = _single(specs[0], qs[0], 0.05, [1e-3, 1e99], [1.0, 0.9], 5000)
    np.testing.assert_array_equal(interior(qb[0]), interior(qg))
    _same_outputs(rb[0], rg)
    specs, qs = _specs(8)
    with Batch2D(specs) as b:
        for k, q in enumerate(qs):
            b.set_q(k, q)
        b.start(0.0, 1.0, [1e-3, 1e99], [1.0, 0.9], 0, METHOD, MTHLIM)
        assert b.advance(10) == 0
        for k in range(len(specs)):
            r = b.result(k)
            assert r["nv"][1] == 0 and r["info"] == 0 and r["tend"] == 0.0
            np.testing.assert_array_equal(interior(b.get_q(k)), interior(qs[k]))


def test_batch_rejects_bad_arguments():
    from classic_amd._lib import ClawGpuError
    specs, _ = _specs(9, sizes=[(10, 10, "euler")], bcs=[(2, 1, 1, 1)])
    with pytest.raises(ClawGpuError):
        Batch2D(specs)  # periodic on one side only
    specs, qs = _specs(9, sizes=[(10, 10, "acoustics")], bcs=[(1, 1, 1, 1)])
    with Batch2D(specs) as b:
        with pytest.raises(ClawGpuError):
            b.start(0.0, 1.0, [1e-3, 1e99], [1.0, 0.9], 10, (1, 2, -1, 0, 0, 0, 0), MTHLIM)
        with pytest.raises(ClawGpuError):
            b.start(0.0, 1.0, [1e-3, 1e99], [0.9, 1.0], 10, METHOD, MTHLIM)  # cflv(2) > cflv(1)
        with pytest.raises(ClawGpuError):
            b.advance(1)  # not started
        q = np.ascontiguousarray(qs[0])
        assert b.lib.clawgpu_batch_set_q(b.h, 1, q.ctypes.data_as(C.POINTER(C.c_double))) != 0
