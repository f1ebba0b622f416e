"""Benchmark: 2D compressible Euler, 4096x4096 cells, 200 steps, FP64.

Workload (the headline configuration): gamma = 1.4, Roe solver with the
Harten-Hyman entropy fix, MC limiter on all 3 waves, unsplit with transverse
corrections (method = 1,2,2), extrapolation BCs, four-quadrant initial data
split at (0.5, 0.5) on [0,1]^2 drawn from numpy.random.default_rng(seed),
CFL 0.9 with variable dt.  A "step" is one accepted time step of the claw2
loop (bc2 + fused step kernel + Courant check) over the whole grid, with q
resident in HBM; the timed region is the device-side claw2 loop of exactly
--steps accepted steps.

Prints one JSON line: value = Gcell-updates/s of the whole run; roofline =
algorithmic HBM bytes (64 B per cell-update: read 4 fields of qold, write 4 of
qnew) per step-kernel launch / its average duration from HIP events on the
library's stream; cpu_baseline = the CPU oracle (the C restatement of the
reference, same arithmetic; OpenMP build on the host's threads) on a
bounded sample of about 10 s.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + %HBM peak; L1 err vs 4x-fine plain run at equal GPU time"
WORKLOAD = "euler2d_4096x4096_roe_efix_mc_unsplit_transverse_200steps"
BYTES_PER_CELL_UPDATE = 64  # (2*meqn + maux) * 8 with meqn = 4, maux = 0
WORKLOAD_A = "acoustics2d_4096x4096_pulse_mc_unsplit_transverse_walls_200steps"
BYTES_PER_CELL_UPDATE_A = 48  # meqn = 3, maux = 0


def measured_ceilings(ndoubles=1 << 27):
    """Ceilings measured on this device: the best of the library's copy kernels
    (1 GiB read + 1 GiB written per copy) and the FP64 FMA throughput of
    independent chains on every SIMD."""
    from classic_amd import _lib

    lib = _lib.load()
    best = None
    for v in range(4):
        ms = C.c_float(0.0)
        _lib.check(lib.clawgpu_copy_bandwidth(ndoubles, v, 5, C.byref(ms)))
        gbs = 2 * ndoubles * 8 / (ms.value * 1e-3) / 1e9
        if best is None or gbs > best[0]:
            best = (gbs, v)
    ms = C.c_float(0.0)
    tf = C.c_double(0.0)
    _lib.check(lib.clawgpu_fp64_fma_peak(20000, C.byref(ms), C.byref(tf)))
    return {"copy_GB_s": best[0], "copy_variant": best[1], "fp64_fma_TFLOP_s": tf.value}


def profile_traffic(name):
    """HBM bytes per step-kernel launch measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a
    plain copy kernel of known size) as recorded in
    profiles/traffic_<name>.json; None until such a measurement exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        try:
            return json.load(open(path)).get("bytes_per_launch")
        except Exception:
            return None
    return None


PORT_NOTE = ("the oracle's C restatement of the reference (same arithmetic, bit for bit with "
             "the reference build in the tests), -O3 OpenMP build: each cell updated in the "
             "serial order by bands of slices")


def cpu_threads():
    """Threads of the CPU baseline: OMP_NUM_THREADS when it is set (16 on the
    GPU box's share of the host), capped by the CPUs this process may run on."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _sample_steps(run, target_s=10.0, cap=64):
    """run(k) -> (accepted steps, seconds) for k steps from the same initial
    state; a one-step probe sizes a sample of about target_s seconds, which
    is what is reported."""
    _, t1 = run(1)
    k = int(min(cap, max(1, round(target_s / max(t1, 1e-3)))))
    return run(k)


def cpu_baseline(mx, my, seed, nsteps=None, solver="euler"):
    """Time the oracle's OpenMP build (kind "port") on claw2 steps of the same
    workload; returns (Gcell-updates/s, seconds, sample, kind, threads)."""
    try:
        from oracle import oracle as O
        from classic_amd import problems as P
        O.lib_omp(1)
    except Exception as e:  # pragma: no cover
        return None, 0.0, f"oracle unavailable: {e}", None, 0
    th = cpu_threads()
    if solver == "acoustics":
        q, dx, dy = P.acoustics_pulse(mx, my)
        p = O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                      method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4], mthbc=(3, 3, 3, 3))
        dt0, name = 0.9 * dx / 2.0, "acoustics"
    else:
        q, dx, dy = P.euler_quadrants(mx, my, seed=seed)
        p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                      method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4, 4], mthbc=(1, 1, 1, 1))
        dt0, name = 0.9 * dx / 3.0, "Euler"

    def run(k):
        t0 = time.perf_counter()
        r = O.claw2(p, q, dx, dy, 0.0, 1e9, [dt0, 1e99], [1.0, 0.9], k, backend="omp", threads=th)
        return int(r.nv[1]), time.perf_counter() - t0

    steps, el = run(nsteps) if nsteps else _sample_steps(run)
    return ((mx * my * steps) / el / 1e9, el,
            f"{mx}x{my} {name}, {steps} claw2 step(s), {th} threads; {PORT_NOTE}", "port", th)


WORKLOAD3 = "acoustics3d_512cubed_vc_layered_superbee_method22_walls_100steps"
WORKLOAD3_SPLIT = "acoustics3d_512cubed_vc_layered_superbee_dimsplit_walls_100steps"
BYTES_PER_CELL_UPDATE3 = 80  # (2*meqn + maux) * 8 with meqn = 4, maux = 2


def cpu_baseline3(n, seed):
    """Time the oracle's OpenMP build (kind "port"; the blocks of each pass of
    the unsplit step on all threads) on unsplit step3 calls of an n^3 sample
    of the 3D workload; returns (Gcell-updates/s, s, sample, kind, threads)."""
    try:
        from oracle import oracle3 as O3
        from classic_amd import problems as P
        O3.lib_omp(1)
    except Exception as e:  # pragma: no cover
        return None, 0.0, f"oracle unavailable: {e}", None, 0
    th = cpu_threads()
    q, aux, _ = P.acoustics3d_pulse(n, seed=seed)
    p = O3.Problem3(n, n, n, mthbc=(3,) * 6, mthlim=(2, 2))
    q = O3.bc3(p, q)
    dt = 0.9 * p.geom.dx / 2.0

    def run(k):
        qq, el = q, 0.0
        for _ in range(k):
            t0 = time.perf_counter()
            qq, _ = O3.step3(p, qq, aux, dt, backend="omp", threads=th)
            el += time.perf_counter() - t0
            qq = O3.bc3(p, qq)
        return k, el

    steps, el = _sample_steps(run)
    return (steps * n ** 3 / el / 1e9, el,
            f"{n}^3 acoustics3d, {steps} unsplit step3 calls, {th} threads; {PORT_NOTE}",
            "port", th)


def bench3(args):
    """3D variable-coefficient acoustics, n^3 cells on [-1,1]^3, two-layer
    medium (Z=1, c=1 below z=0; Z=6, c=2 above), Gaussian pressure pulse at a
    seeded position, walls on all faces, superbee on both waves, unsplit
    method (2, 22), CFL 0.9: --steps accepted claw3 steps, q and aux
    resident in HBM."""
    from classic_amd import problems as P
    from classic_amd.gpu3 import Grid3D

    n = args.n3
    q0, aux, _ = P.acoustics3d_pulse(n, seed=args.seed)
    method, mthlim, mthbc = (1, 2, args.m3, 0, 0, 0, 2), (2, 2), (3, 3, 3, 3, 3, 3)
    g = Grid3D(n, n, n)
    if args.tile >= 0 or args.tile_order >= 0:
        tile, order = g.get_tile()
        g.set_tile(args.tile if args.tile >= 0 else tile,
                   args.tile_order if args.tile_order >= 0 else order)
    tile_used = g.get_tile()
    g.set_aux(aux)
    g.set_q(q0)
    dx = g.dx[0]
    # warmup: an untimed claw3 run of the same problem on a second grid,
    # after q0 is resident in the timed grid, so the timed run starts on a
    # GPU already at its steady clock (no upload in between)
    gw = None
    if args.warmup > 0:
        gw = Grid3D(n, n, n)
        gw.set_tile(*tile_used)
        gw.set_aux(aux)
        gw.set_q(q0)
        gw.claw3(0.0, 1e30, [0.9 * dx / 2.0, 1e99], [1.0, 0.9], args.warmup, method, mthlim,
                 mthbc)
    del aux, q0
    t0 = time.perf_counter()
    r = g.claw3(0.0, 1e30, [0.9 * dx / 2.0, 1e99], [1.0, 0.9], args.steps, method, mthlim, mthbc)
    wall = time.perf_counter() - t0
    if gw is not None:
        gw.close()
    steps = int(r.nv[1])
    if steps != args.steps:
        raise SystemExit(f"claw3 took {steps} steps, expected {args.steps} (info={r.info})")
    cells = n ** 3
    value = cells * steps / wall / 1e9
    ksteps = max(5, min(args.steps, 20))
    total_ms, kern_ms, _ = g.time_steps(float(r.dtv[4]), ksteps, method, mthlim, mthbc)
    g.close()
    kavg_s = kern_ms / ksteps / 1e3
    achieved = cells * BYTES_PER_CELL_UPDATE3 / kavg_s / 1e9
    ceil = measured_ceilings()
    cpu = None
    if not args.no_cpu:
        v, el, sample, kind, th = cpu_baseline3(args.cpu_n3, args.seed)
        if v is not None:
            cpu = {"value": v, "unit": "Gcell-updates/s", "cores": th, "kind": kind,
                   "nproc": os.cpu_count(), "sample": f"{sample} ({el:.1f} s)"}
    return {
        "metric": METRIC, "value": value, "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian pulse in a two-layer medium)",
        "config": {"workload": (WORKLOAD3 if args.m3 == 22 else WORKLOAD3_SPLIT) if n == 512
                   else f"acoustics3d_{n}cubed_method{args.m3}", "mx": n,
                   "my": n, "mz": n, "seed": args.seed, "method": list(method),
                   "mthlim": list(mthlim), "mthbc": list(mthbc), "cfl_desired": 0.9,
                   "cfl_max": 1.0},
        "tile_shape": tile_used[0], "tile_order": tile_used[1],
        "hbm_frac_of_peak": value * BYTES_PER_CELL_UPDATE3 / HBM_PEAK_GBS,
        "kernel_ms_avg": kavg_s * 1e3,
        "step_overhead_ms_avg": (total_ms - kern_ms) / ksteps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (profile_traffic("acoustics3d512")
                                 if n == 512 and args.m3 == 22 else None)},
        "measured_ceilings": ceil,
        "frac_of_measured_copy": achieved / ceil["copy_GB_s"],
        "cpu_baseline": cpu,
    }


WORKLOAD_B = "batch2d_256problems_sides_loguniform_128_1024_euler_acoustics_mc_unsplit_transverse"


def cpu_baseline_batch(specs, min_cells=200_000_000):
    """Time the oracle's OpenMP build (kind "port") on one claw2 step of each
    batch problem in turn until min_cells cell updates are done; returns
    (Gcell-updates/s, s, sample, kind, threads)."""
    try:
        from oracle import oracle as O
        from classic_amd import problems as P
        O.lib_omp(1)
    except Exception as e:  # pragma: no cover
        return None, 0.0, f"oracle unavailable: {e}", None, 0
    th = cpu_threads()
    done, el, used = 0, 0.0, 0
    for s in specs:
        q, dx, dy = P.batch_problem_q(s)
        if s["solver"] == "euler":
            p = O.Problem(4, 3, s["mx"], s["my"], O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                          method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4, 4], mthbc=s["mthbc"])
        else:
            p = O.Problem(3, 2, s["mx"], s["my"], O.SOLVER_ACOUSTICS, O.solver_params(*s["params"]),
                          method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4], mthbc=s["mthbc"])
        t0 = time.perf_counter()
        r = O.claw2(p, q, dx, dy, 0.0, 1e9, [1e-4, 1e99], [1.0, 0.9], 1, backend="omp", threads=th)
        el += time.perf_counter() - t0
        done += s["mx"] * s["my"] * int(r.nv[1])
        used += 1
        if done >= min_cells or el >= 20.0:
            break
    return (done / el / 1e9, el, (f"first {used} problems of the batch, 1 claw2 step each "
                                  f"({done} cell updates), {th} threads; {PORT_NOTE}"),
            "port", th)


def bench_batch(args):
    """256 independent 2D problems (sides log-uniform in [128, 1024], Euler
    four-quadrant or random-coefficient acoustics), each on its own claw2
    loop with its own dt, all advanced by one persistent kernel per batched
    iteration, replayed from a captured HIP graph; the schedule comes from
    the autotuner.  value = accepted cell updates per second over --steps
    batched iterations."""
    from classic_amd import problems as P
    from classic_amd.batch import from_specs

    specs = P.batch_problem_specs(args.nprob, seed=args.seed)
    arith = 1 if args.arith == "fast" else 0
    b, qs = from_specs(specs, arith=arith)
    del qs
    method, mthlim = (1, 2, 2, 0, 0, 0, 0), (4, 4, 4)
    b.start(0.0, 1e30, [1e-4, 1e99], [1.0, 0.9], 10 ** 9, method, mthlim)
    cands = b.tune(reps=2)
    sched = b.schedule()
    if args.warmup > 0:
        b.advance(args.warmup)
    cells = np.array([s["mx"] * s["my"] for s in specs], dtype=np.int64)
    meqn = np.array([4 if s["solver"] == "euler" else 3 for s in specs], dtype=np.int64)
    nv0 = np.array([b.result(k)["nv"][1] for k in range(len(specs))], dtype=np.int64)
    t0 = time.perf_counter()
    b.advance(args.steps)
    wall = time.perf_counter() - t0
    res = [b.result(k) for k in range(len(specs))]
    nv1 = np.array([r["nv"][1] for r in res], dtype=np.int64)
    # a problem stops early only on an error return (info -5: NaN wave speed,
    # e.g. a Roe state with negative pressure); it then adds no updates
    stopped = {}
    for r in res:
        if r["info"] != 0:
            stopped[str(r["info"])] = stopped.get(str(r["info"]), 0) + 1
    updates = int(np.sum((nv1 - nv0) * cells))
    value = updates / wall / 1e9
    ksteps = max(5, min(args.steps, 20))
    total_ms, kern_ms = b.time_iterations(ksteps)
    b.close()
    kavg_s = kern_ms / ksteps / 1e3
    bytes_iter = int(np.sum(2 * meqn * 8 * cells))
    achieved = bytes_iter / kavg_s / 1e9
    cpu = None
    if not args.no_cpu:
        v, el, sample, kind, th = cpu_baseline_batch(specs)
        if v is not None:
            cpu = {"value": v, "unit": "Gcell-updates/s", "cores": th, "kind": kind,
                   "nproc": os.cpu_count(), "sample": f"{sample} ({el:.1f} s)"}
    best = min(cands, key=lambda c: c[4])
    worst = max(cands, key=lambda c: c[4])
    return {
        "metric": METRIC,
        "value": value,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded batch: four-quadrant Euler and acoustics pulses)",
        "config": {"workload": WORKLOAD_B, "nprob": len(specs), "seed": args.seed,
                   "cells_total": int(cells.sum()),
                   "euler_problems": int(sum(s["solver"] == "euler" for s in specs)),
                   "method": list(method), "mthlim": list(mthlim), "cfl_desired": 0.9,
                   "cfl_max": 1.0, "schedule": sched},
        "arith": args.arith,
        "accepted_updates": updates,
        "problems_stopped_by_info": stopped,
        "kernel_ms_avg": kavg_s * 1e3,
        "iteration_ms_avg_eager": total_ms / ksteps,
        "autotuner": {"candidates": len(cands), "best_ms": best[4], "worst_ms": worst[4]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (profile_traffic("batch2d256")
                                 if len(specs) == 256 and args.seed == 0 else None)},
        "cpu_baseline": cpu,
    }


WORKLOAD_L = "learned2d_1024x1024_euler_mlp_limiter_unsplit_transverse_200steps"
MLP_FLOP_PER_WAVE = 2 * (5 * 32 + 32 * 32 + 32)  # 5 -> 32 -> 32 -> 1 multiply-adds
MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_learned(args):
    """The learned-limiter run: 2D Euler at n^2 (1024^2 by default) with
    limiter 6 on all three waves (features pass + BF16-MFMA network + step
    kernel reading the phi tables every step), network weights of the
    5-32-32-1 architecture shipped with the package (trained by
    classic_amd.learned on 4x-coarsened plain 4096^2 runs).  value = Gcell-updates/s of the learned run; the roofline is
    that of the fused pre-pass kernel a learned step runs (features in FP64,
    then the network on the matrix cores for the waves that need it):
    2,432 network flop per evaluated wave (nonzero, theta > 0) / its average
    launch time.
    l1_vs_fine_4x: the metric's L1 part -- learned at n^2, MC at n^2 and MC
    at (2n)^2 against the plain (4n)^2 run on held-out seeds 200 and 201 to
    t = 0.1, each with its GPU time, and the learned run's error and time
    over MC at (2n)^2's (the gap to the target where the ratio exceeds 1)."""
    from classic_amd import SOLVER_EULER
    from classic_amd import learned as LL
    from classic_amd import problems as P

    n = args.n
    arith = 1 if args.arith == "fast" else 0
    q0, dx, dy = P.euler_quadrants(n, n, seed=args.seed)
    method, mthbc = (1, 2, 2, 0, 0, 0, 0), (1, 1, 1, 1)
    mth_l, mth_mc = [6, 6, 6], [4, 4, 4]
    dt0 = 0.9 * dx / 3.0
    g = LL.LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith)
    g.set_params(LL.trained_params())
    g.set_q(q0)
    if args.warmup > 0:
        g.time_steps(dx, dy, 0.25 * dx, method, mth_l, mthbc, args.warmup)
    runs = {}
    for name, lim in (("mc", mth_mc), ("learned", mth_l)):
        g.set_q(q0)
        t0 = time.perf_counter()
        out = g.claw2(dx, dy, 0.0, 1e30, [dt0, 1e99], [1.0, 0.9], args.steps, method, lim, mthbc)
        wall = time.perf_counter() - t0
        steps = int(out["nv"][1])
        if steps != args.steps:
            raise SystemExit(f"claw2 ({name}) took {steps} steps, expected {args.steps}")
        runs[name] = (wall, float(out["dtv"][4]))
    wall, dt_last = runs["learned"]
    value = n * n * args.steps / wall / 1e9
    reps = 20
    fused_ms, infer_ms = g.time_kernels(dx, dy, dt_last, reps)
    ksteps = 20
    _, step_ms, _ = g.time_steps(dx, dy, dt_last, method, mth_l, mthbc, ksteps)
    _, step_mc_ms, _ = g.time_steps(dx, dy, dt_last, method, mth_mc, mthbc, ksteps)
    waves = 2 * g.nwaves()
    # the waves the fused pre-pass runs through the network at this state:
    # nonzero, with a nonzero strength feature and theta > 0 (every other
    # wave takes MC(theta) or 0 without the network); the feature pass of
    # time_kernels left the feature and theta tables of the same q and dt
    feat = g.read(0).reshape(-1, 8)
    theta = g.read(3)
    evaluated = int(np.count_nonzero((feat[:, 3] > 0) & (theta > 0)))
    g.close()
    # second half of the metric: L1 error against the 4x-fine plain run, with
    # the GPU time of each run (wall time of its claw2 loop, which waits for
    # the device), on held-out seed 200 to t = 0.1
    l1 = {"tend": 0.1, "reference": f"MC at {4 * n}^2, 4x4 block-averaged", "seeds": {}}
    for ev in LL.evaluate(LL.trained_params(), (200, 201), n, 0.1):
        l1["seeds"][str(ev["seed"])] = {
            f"learned_{n}": {"l1": ev["l1_learned"], "gpu_s": ev["t_learned"],
                             "steps": ev["steps_learned"]},
            f"mc_{n}": {"l1": ev["l1_mc"], "gpu_s": ev["t_mc"], "steps": ev["steps_mc"]},
            f"mc_{2 * n}": {"l1": ev["l1_mc2x"], "gpu_s": ev["t_mc2x"], "steps": ev["steps_mc2x"]},
            # > 1: the learned run's error is above MC at twice the resolution
            "l1_learned_over_mc2x": ev["l1_learned"] / ev["l1_mc2x"],
            "gpu_time_learned_over_mc2x": ev["t_learned"] / ev["t_mc2x"]}
    flops = waves * MLP_FLOP_PER_WAVE
    achieved = evaluated * MLP_FLOP_PER_WAVE / (fused_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        # the MC run of the same grid on the reference's CPU path (it has no
        # learned limiter)
        v, el, sample, kind, th = cpu_baseline(n, n, args.seed, None, "euler")
        if v is not None:
            cpu = {"value": v, "unit": "Gcell-updates/s", "cores": th, "kind": kind,
                   "nproc": os.cpu_count(), "sample": f"{sample}, MC limiter ({el:.1f} s)"}
    return {
        "metric": METRIC,
        "value": value,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, bf16 network with f32 accumulation",
        "data": "synthetic (seeded four-quadrant Euler initial data; the shipped trained network)",
        "config": {"workload": WORKLOAD_L, "mx": n, "my": n, "seed": args.seed,
                   "method": list(method), "mthlim": mth_l, "mthbc": list(mthbc),
                   "mlp": "5-32-32-1 GELU", "cfl_desired": 0.9, "cfl_max": 1.0},
        "arith": args.arith,
        "mc_ms_per_step": runs["mc"][0] / args.steps * 1e3,
        "kernel_ms": {"fused_prepass": fused_ms, "mlp_inference_alone": infer_ms,
                      "step_with_learned": step_ms / ksteps, "step_mc": step_mc_ms / ksteps},
        "l1_vs_fine_4x": l1,
        "mlp_waves_per_launch": waves,
        "mlp_waves_evaluated": evaluated,
        "mlp_inference_alone_tflops": flops / (infer_ms * 1e-3) / 1e12,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MFMA_BF16_PEAK_TFLOPS,
                     "traffic": (profile_traffic("learned1024")
                                 if n == 1024 and args.arith == "fast" else None)},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload",
                    choices=["euler2d", "acoustics2d", "acoustics3d", "batch2d", "learned2d"],
                    default="euler2d",
                    help="euler2d: the headline 4096^2 Euler run; acoustics2d: the 200^2 "
                         "acoustics pulse problem at 4096^2; acoustics3d: the 512^3 "
                         "variable-coefficient acoustics run; batch2d: 256 uneven 2D problems "
                         "in one persistent kernel; learned2d: 1024^2 Euler with the MLP "
                         "limiter")
    ap.add_argument("--nprob", type=int, default=256)
    ap.add_argument("--n3", type=int, default=512)
    ap.add_argument("--cpu-n3", type=int, default=160)
    ap.add_argument("--tile", type=int, default=-1,
                    help="acoustics3d: tile shape of the unsplit kernel (0..4 tile kernels, "
                         "5, 6 z-marching kernel; -1: the library's default)")
    ap.add_argument("--m3", type=int, default=22, choices=[22, -1],
                    help="acoustics3d: method(3), 22 (unsplit, both corrections; the headline) "
                         "or -1 (dimensional splitting, dimsp3)")
    ap.add_argument("--tile-order", type=int, default=-1,
                    help="acoustics3d: tile order of the unsplit kernel (bit 0 XCD-contiguous, "
                         "bit 1 z fastest; -1: the library's default)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="accepted steps to time (default 200 for euler2d, 100 for acoustics3d)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warmup steps (default 100 for euler2d/acoustics2d, 10 for "
                         "acoustics3d, each run as a claw run of the same problem on a second "
                         "grid right before the timed run; 5 otherwise)")
    ap.add_argument("--n", type=int, default=None,
                    help="grid side (default 4096; 1024 for learned2d)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the CPU baseline sample (0: sized to about 10 s)")
    ap.add_argument("--arith", choices=["fast", "exact"], default="fast",
                    help="fast: FMA + reciprocal reuse, held to 1e-12 of the reference by the "
                         "tests; exact: the reference's rounding bit for bit")
    ap.add_argument("--seg-rows", type=int, default=0,
                    help="rows per wavefront segment of the step kernel (0: the library's "
                         "per-grid choice)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of a known size (for the rocprof "
                         "FETCH_SIZE/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the single-grid step is not sharded across GPUs.", file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = 100 if args.workload == "acoustics3d" else 200
    if args.warmup is None:
        args.warmup = {"euler2d": 100, "acoustics2d": 100, "acoustics3d": 10}.get(args.workload, 5)
    if args.workload == "acoustics3d":
        print(json.dumps(bench3(args)), flush=True)
        return
    if args.workload == "batch2d":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.n is None:
        args.n = 1024 if args.workload == "learned2d" else 4096
    if args.workload == "learned2d":
        print(json.dumps(bench_learned(args)), flush=True)
        return

    from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
    from classic_amd import problems as P

    n = args.n
    acoustics = args.workload == "acoustics2d"
    arith = 1 if args.arith == "fast" else 0
    if acoustics:
        # the pulse problem of the 200^2 acoustics configuration at n^2
        q0, dx, dy = P.acoustics_pulse(n, n)
        method, mthlim, mthbc = (1, 2, 2, 0, 0, 0, 0), [4, 4], (3, 3, 3, 3)
        g = gpu.Grid2D(SOLVER_ACOUSTICS, n, n, [1.0, 4.0], arith=arith)
        dt0 = 0.9 * dx / 2.0  # c = 2
        workload, bytes_cell, traffic_name = WORKLOAD_A, BYTES_PER_CELL_UPDATE_A, "acoustics4096"
        data = "synthetic (the acoustics pulse problem's initial data at this size)"
    else:
        q0, dx, dy = P.euler_quadrants(n, n, seed=args.seed)
        method, mthlim, mthbc = (1, 2, 2, 0, 0, 0, 0), [4, 4, 4], (1, 1, 1, 1)
        g = gpu.Grid2D(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith)
        # first dt from the CFL-desired rule of the initial data: the retake
        # logic corrects it on the first step if it is too large
        dt0 = 0.9 * dx / 3.0
        workload, bytes_cell, traffic_name = WORKLOAD, BYTES_PER_CELL_UPDATE, "euler4096"
        data = "synthetic (seeded four-quadrant Euler initial data)"
    g.set_segment_rows(args.seg_rows)
    g.set_q(q0)
    # warmup: an untimed claw2 run of the same problem on a second grid,
    # started after q0 is resident in the timed grid, so that the timed run
    # follows it with no idle gap (this FP64 load takes the GPU about 50 ms
    # of continuous work to reach its steady clock; after an idle gap the
    # first steps run up to 30 % slower)
    gw = None
    if args.warmup > 0:
        gw = gpu.Grid2D(SOLVER_ACOUSTICS if acoustics else SOLVER_EULER, n, n,
                        [1.0, 4.0] if acoustics else [1.4, 1.0], arith=arith)
        gw.set_segment_rows(args.seg_rows)
        gw.set_q(q0)
        gw.claw2(dx, dy, 0.0, 1e30, [dt0, 1e99], [1.0, 0.9], args.warmup, method, mthlim, mthbc)

    t0 = time.perf_counter()
    out = g.claw2(dx, dy, 0.0, 1e30, [dt0, 1e99], [1.0, 0.9], args.steps, method, mthlim, mthbc)
    wall = time.perf_counter() - t0
    if gw is not None:
        gw.close()
    steps = int(out["nv"][1])
    if steps != args.steps:
        raise SystemExit(f"claw2 took {steps} steps, expected {args.steps} (info={out['info']})")
    cells = n * n
    value = cells * steps / wall / 1e9

    # per-launch kernel time at the final dt (same data, continuing the run)
    dt_last = float(out["dtv"][4])
    ksteps = max(10, min(args.steps, 50))
    total_ms, kern_ms, _ = g.time_steps(dx, dy, dt_last, method, mthlim, mthbc, ksteps)
    kavg_s = kern_ms / ksteps / 1e3
    achieved = cells * bytes_cell / kavg_s / 1e9
    seg_rows = g.segment_rows()
    # the same kernel time with the other arithmetic, for reference
    g.set_arith(1 - arith)
    _, kern_other, _ = g.time_steps(dx, dy, dt_last, method, mthlim, mthbc, ksteps)
    g.close()
    ceil = measured_ceilings()
    copy_cal = None
    if args.copy_calibration:
        from classic_amd import _lib

        ms = C.c_float(0.0)
        ncal = 1 << 27  # 1 GiB per buffer
        _lib.check(_lib.load().clawgpu_copy_calibration(ncal, 3, C.byref(ms)))
        copy_cal = {"bytes_each_way": ncal * 8, "ms": ms.value,
                    "GB_s": 2 * ncal * 8 / (ms.value * 1e-3) / 1e9}

    cpu = None
    if not args.no_cpu:
        v, el, sample, kind, th = cpu_baseline(n, n, args.seed, args.cpu_steps or None,
                                               "acoustics" if acoustics else "euler")
        if v is not None:
            cpu = {"value": v, "unit": "Gcell-updates/s", "cores": th, "kind": kind,
                   "nproc": os.cpu_count(), "sample": f"{sample} ({el:.1f} s)"}

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": data,
        "config": {"workload": workload, "mx": n, "my": n, "seed": args.seed,
                   "method": list(method), "mthlim": mthlim, "mthbc": list(mthbc),
                   "cfl_desired": 0.9, "cfl_max": 1.0},
        "hbm_frac_of_peak": value * bytes_cell / HBM_PEAK_GBS,
        "arith": args.arith,
        "kernel_ms_avg": kavg_s * 1e3,
        "segment_rows": seg_rows,
        "kernel_ms_avg_other_arith": kern_other / ksteps,
        "step_overhead_ms_avg": (total_ms - kern_ms) / ksteps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (profile_traffic(traffic_name)
                                 if n == 4096 and args.arith == "fast" else None)},
        "measured_ceilings": ceil,
        "frac_of_measured_copy": achieved / ceil["copy_GB_s"],
        "cpu_baseline": cpu,
    }
    if copy_cal:
        rec["copy_calibration"] = copy_cal
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
