"""Learned flux limiter: device front end, training and evaluation.

The network (5 -> 32 -> 32 -> 1, GELU) replaces phi(theta) of the limiter as
limiter method 6 (see include/clawgpu.h).  Everything numerical runs in the
library's HIP kernels: the feature pre-pass, the BF16 MFMA inference with the
TVD clip and MC fallback, the FP32 forward/backward that produces parameter
gradients, and the one-step loss against a target state.  This module only
orchestrates: it generates training data with the plain method (4x-coarsened
fine runs), applies Adam updates to the 1281 parameters on the host, and
measures L1 errors against the fine solution.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import problems as P
from ._lib import SOLVER_EULER, check
from .gpu import Grid2D, _dp, _ip, _ivec

NPARAMS = 1281
H = 32
METHOD = (1, 2, 2, 0, 0, 0, 0)
MTHBC = (1, 1, 1, 1)


def init_params(seed=0):
    """He-style initialisation, output bias 1 (the MC value of a smooth
    wave), float32 in the library's order W1, b1, W2, b2, W3, b3."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((5, H)) * np.sqrt(2.0 / 5)
    W2 = rng.standard_normal((H, H)) * np.sqrt(2.0 / H)
    W3 = rng.standard_normal(H) * np.sqrt(1.0 / H) * 0.1
    parts = [W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3, np.array([1.0])]
    return np.concatenate(parts).astype(np.float32)


class LearnedGrid(Grid2D):
    """A Grid2D with the learned-limiter calls of the C-ABI."""

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(self.lib.clawgpu_learned_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                  p.size))

    def features(self, dx, dy, dt):
        check(self.lib.clawgpu_learned_features(self.h, dx, dy, dt))

    def phi(self, mode):
        check(self.lib.clawgpu_learned_phi(self.h, mode))

    def time_kernels(self, dx, dy, dt, reps):
        """Average ms per launch of the fused pre-pass of a limiter-6 step
        (features and network in one kernel) and of the stand-alone MFMA
        inference over stored features, on the current q (HIP events on the
        library's stream)."""
        f, i = C.c_float(0.0), C.c_float(0.0)
        check(self.lib.clawgpu_learned_time(self.h, dx, dy, dt, reps, C.byref(f), C.byref(i)))
        return f.value, i.value

    def nwaves(self):
        n = C.c_long(0)
        check(self.lib.clawgpu_learned_table_size(self.h, C.byref(n)))
        return n.value

    def read(self, what):
        nw = self.nwaves()
        spec = {0: (np.float32, (2 * nw, 8)), 1: (np.float64, (nw,)), 2: (np.float64, (nw,)),
                3: (np.float64, (2 * nw,)), 4: (np.float32, (2 * nw, self.meqn)),
                5: (np.float32, (2 * nw,))}[what]
        out = np.empty(spec[1], dtype=spec[0])
        check(self.lib.clawgpu_learned_read(self.h, what, out.ctypes.data_as(C.c_void_p),
                                            out.nbytes))
        return out

    def write(self, what, arr):
        arr = np.ascontiguousarray(arr)
        check(self.lib.clawgpu_learned_write(self.h, what, arr.ctypes.data_as(C.c_void_p),
                                             arr.nbytes))

    def set_target(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        check(self.lib.clawgpu_learned_set_target(self.h, _dp(q)))

    def grad(self, dx, dy, dt, mode, method=METHOD, mthbc=MTHBC):
        g = np.zeros(NPARAMS, dtype=np.float32)
        loss = C.c_double(0.0)
        check(self.lib.clawgpu_learned_grad(self.h, dx, dy, dt, _ip(_ivec(method, 7)),
                                            _ip(_ivec(mthbc, 4)), mode,
                                            g.ctypes.data_as(C.POINTER(C.c_float)),
                                            C.byref(loss)))
        return g.astype(np.float64), loss.value


def coarsen(q, f, mbc=2):
    """Block-average the interior of q (reference layout) by f x f; ghost
    cells of the result are zero."""
    qi = q[mbc:-mbc, mbc:-mbc]
    my, mx, m = qi.shape
    c = qi.reshape(my // f, f, mx // f, f, m).mean(axis=(1, 3))
    out = np.zeros((my // f + 2 * mbc, mx // f + 2 * mbc, m))
    out[mbc:-mbc, mbc:-mbc] = c
    return out


def fine_snapshots(n_coarse, seed, nsnap, factor=4, cfl=0.8, arith=1, span=None):
    """Run the plain method on the fine grid (factor * n_coarse)^2 with a
    fixed dt (`cfl` / 1.6 times the Courant limit of the initial data) and
    return coarsened snapshots [(q_t, q_t+dtc)] separated by `factor` fine
    steps (one coarse step of dtc = factor * dt_fine), plus the coarse dx
    and dtc.  Without `span` the pairs start after 8 coarse steps and are 3
    coarse steps apart (early times); with `span` (a physical time) they are
    spread evenly over (0, span]."""
    nf = n_coarse * factor
    q0, dxf, dyf = P.euler_quadrants(nf, nf, seed=seed)
    # max wave speed of the initial data
    rho, u, v = q0[2:-2, 2:-2, 0], q0[2:-2, 2:-2, 1] / q0[2:-2, 2:-2, 0], q0[2:-2, 2:-2, 2] / q0[2:-2, 2:-2, 0]
    p = 0.4 * (q0[2:-2, 2:-2, 3] - 0.5 * rho * (u * u + v * v))
    smax = (np.sqrt(u * u + v * v) + np.sqrt(1.4 * p / rho)).max()
    dtf = cfl * dxf / (smax * 1.6)
    warm, gap = 8, 2
    if span is not None:
        per = max(1, int(round(span / (nsnap * dtf * factor))))
        warm, gap = per, per - 1
    out = []
    with Grid2D(SOLVER_EULER, nf, nf, [1.4, 1.0], arith=arith) as g:
        g.set_q(q0)
        # let the waves develop before sampling
        g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, warm * factor)
        for k in range(nsnap):
            a = coarsen(g.get_q(), factor)
            g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, factor)
            b = coarsen(g.get_q(), factor)
            out.append((a, b))
            if k + 1 < nsnap and gap > 0:
                g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, gap * factor)
    return out, dxf * factor, dtf * factor


def fine_pairs(n_coarse, seed, nsnap, courant, span, factor=4, arith=1):
    """As fine_snapshots, with the fine dt set from the measured Courant
    number: after each stretch of fine steps dt is rescaled so the next one
    runs at `courant`, and the pairs start at span * k / nsnap, k = 1..nsnap.
    A pair whose own steps exceed Courant number 1 raises.  Returns
    [(q_t, q_t+dtc, dx_coarse, dtc)]."""
    nf = n_coarse * factor
    q0, dxf, dyf = P.euler_quadrants(nf, nf, seed=seed)
    rho = q0[2:-2, 2:-2, 0]
    u, v = q0[2:-2, 2:-2, 1] / rho, q0[2:-2, 2:-2, 2] / rho
    p = 0.4 * (q0[2:-2, 2:-2, 3] - 0.5 * rho * (u * u + v * v))
    dtf = 0.5 * dxf / (np.sqrt(u * u + v * v) + np.sqrt(1.4 * p / rho)).max()
    lim = [4, 4, 4]
    out, t = [], 0.0
    with Grid2D(SOLVER_EULER, nf, nf, [1.4, 1.0], arith=arith) as g:
        g.set_q(q0)
        for k in range(1, nsnap + 1):
            # advance to the pair's start, rescaling dt after each stretch
            while True:
                n = int((span * k / nsnap - t) / dtf)
                if n <= 0:
                    break
                n = min(n, 8 * factor)
                cfl = g.time_steps(dxf, dyf, dtf, METHOD, lim, MTHBC, n)[2]
                if not cfl <= 1.0:
                    raise RuntimeError(f"fine run exceeded Courant number 1 ({cfl})")
                t += n * dtf
                dtf *= courant / cfl
            a = coarsen(g.get_q(), factor)
            cfl = g.time_steps(dxf, dyf, dtf, METHOD, lim, MTHBC, factor)[2]
            if not cfl <= 1.0:
                raise RuntimeError(f"fine run exceeded Courant number 1 ({cfl})")
            t += factor * dtf
            out.append((a, coarsen(g.get_q(), factor), dxf * factor, dtf * factor))
    return out


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)


def train(n_coarse=256, seeds=(11, 12, 13), nsnap=4, pre_iters=300, iters=300, lr=2e-3,
          seed=0, log=None, cfl=0.8, span=None, courant=None, pretarget="mc", val_seed=None,
          val_every=25, val_tend=0.1, loss="mse", lr2=None, init=None, es_iters=0,
          es_pairs=6, es_sigma=0.02, es_lr=0.01, es_tend=0.1, es_from=None):
    """Train the network: (1) regression onto a classical limiter (pretarget
    "mc" or "superbee": the network starts as that limiter), (2) the one-step
    loss of the unsplit step against 4x-coarsened fine solutions
    (fine_snapshots with cfl, span; fine_pairs when courant is given),
    squared ("mse") or absolute ("l1") error per cell, learning rate lr2
    (lr / 2 by default).  init: starting parameters in place of
    init_params(seed) (for instance a stage-1 result with pre_iters 0).
    Stage 3 (es_iters > 0) minimises the multi-step error itself: the mean
    L1 error of n_coarse^2 runs of the training seeds to es_tend against
    their 4x-fine plain runs, by antithetic evolution strategies (es_pairs
    Gaussian directions of scale es_sigma per iteration, Adam step es_lr)
    over the parameters from offset es_from on (default: the second hidden
    bias and the output layer); the parameters with the smallest objective
    are kept.  Returns (params, history)."""
    params, hist = _train12(n_coarse, seeds, nsnap, pre_iters, iters, lr, seed, log, cfl, span,
                            courant, pretarget, val_seed, val_every, val_tend, loss, lr2, init)
    if es_iters > 0:
        params = _train_es(params.astype(np.float64), n_coarse, seeds, es_iters, es_pairs,
                           es_sigma, es_lr, es_tend, seed, hist, log,
                           _ES_FROM if es_from is None else es_from).astype(np.float32)
    return params, hist


# offset of b2 (the second hidden layer's bias) in the parameter vector:
# W1 (5 x 32), b1, W2 (32 x 32), then b2, W3, b3
_ES_FROM = 5 * H + H + H * H


def _train_es(params, n, seeds, iters, pairs, sigma, lr, tend, seed, hist, log, start):
    """Stage 3 of train: antithetic evolution strategies on the multi-step
    L1 error (Salimans et al. 2017, with an Adam step)."""
    refs = {}
    for s in seeds:
        qf, _, _ = run(4 * n, s, tend, 4)
        refs[s] = coarsen(qf, 4)
        del qf

    def objective(p):
        return float(np.mean([l1_error(run(n, s, tend, 6, p.astype(np.float32))[0], refs[s])
                              for s in seeds]))

    best, fs = es_minimise(objective, params, start, iters, pairs, sigma, lr,
                           np.random.default_rng(seed + 7919), log)
    hist["es"] = fs
    return best


def es_minimise(objective, params, start, iters, pairs, sigma, lr, rng, log=None):
    """Antithetic evolution strategies over params[start:]: per iteration,
    the gradient estimate sum_k (f(p + sigma e_k) - f(p - sigma e_k)) e_k /
    (2 pairs sigma) over Gaussian directions e_k feeds an Adam step.  Returns
    the parameters with the smallest objective seen and the objective after
    each iteration (the starting value first)."""
    params = np.array(params, dtype=np.float64)
    idx = np.arange(start, len(params))
    opt = Adam(len(idx), lr=lr)
    sub = params[idx].copy()
    best = (objective(params), params.copy())
    fs = [best[0]]
    for it in range(iters):
        g = np.zeros(len(idx))
        for _ in range(pairs):
            e = rng.standard_normal(len(idx))
            pp, pm = params.copy(), params.copy()
            pp[idx] += sigma * e
            pm[idx] -= sigma * e
            g += (objective(pp) - objective(pm)) * e
        sub = opt.step(sub, g / (2 * pairs * sigma))
        params[idx] = sub
        f = objective(params)
        fs.append(f)
        if f < best[0]:
            best = (f, params.copy())
        if log:
            log(f"es iter {it + 1} objective {f:.5e} (best {best[0]:.5e})")
    return best[1], fs


def _train12(n_coarse, seeds, nsnap, pre_iters, iters, lr, seed, log, cfl, span, courant,
             pretarget, val_seed, val_every, val_tend, loss, lr2, init):
    """Stages 1 and 2 of train.  With val_seed, the network kept is the one
    (after stage 1 or at any multiple of val_every in stage 2) with the
    smallest multi-step L1 error of a run of that seed to val_tend against
    its 4x-fine plain run.  Returns (params, history)."""
    data = []
    for s in (seeds if pre_iters > 0 or iters > 0 else ()):
        if courant is not None:
            data += fine_pairs(n_coarse, s, nsnap, courant, span if span else 0.1)
        else:
            snaps, dx, dtc = fine_snapshots(n_coarse, s, nsnap, cfl=cfl, span=span)
            data += [(a, b, dx, dtc) for a, b in snaps]
    params = (init_params(seed) if init is None else np.asarray(init)).astype(np.float64)
    hist = {"pre": [], "step": [], "step_mc": [], "val": [], "val_mc": None}
    premode = {"mc": 0, "superbee": 2}[pretarget]
    ref = None
    if val_seed is not None:
        qf, _, _ = run(4 * n_coarse, val_seed, val_tend, 4)
        ref = coarsen(qf, 4)
        del qf
        hist["val_mc"] = l1_error(run(n_coarse, val_seed, val_tend, 4)[0], ref)
    best = [np.inf, None]

    def validate(it, p):
        if ref is None:
            return
        l1 = l1_error(run(n_coarse, val_seed, val_tend, 6, p.astype(np.float32))[0], ref)
        hist["val"].append((it, l1))
        if l1 < best[0]:
            best[:] = [l1, p.copy()]

    t0 = time.time()
    with LearnedGrid(SOLVER_EULER, n_coarse, n_coarse, [1.4, 1.0], arith=1) as g:
        # stage 1: regression onto the classical limiter over the features
        # of all snapshots (loss per live wave)
        opt = Adam(NPARAMS, lr=lr)
        nlive = {}
        for it in range(pre_iters):
            k = it % len(data)
            a, b, dx, dtc = data[k]
            g.set_q(a)
            g.bc2(MTHBC)
            g.features(dx, dx, dtc)
            if k not in nlive:
                nlive[k] = max(1, int((g.read(0)[:, 3] > 0).sum()))
            g.set_params(params)
            gr, lv = g.grad(dx, dx, dtc, premode)
            params = opt.step(params, gr / nlive[k])
            hist["pre"].append(lv / nlive[k])
        validate(0, params)
        # baseline one-step loss of MC on every snapshot
        for a, b, dx, dtc in data:
            hist["step_mc"].append(one_step_loss(g, a, b, dx, dtc, None))
        # stage 2: one-step loss
        opt = Adam(NPARAMS, lr=lr * 0.5 if lr2 is None else lr2)
        mode = {"mse": 1, "l1": 3}[loss]
        for it in range(iters):
            a, b, dx, dtc = data[it % len(data)]
            g.set_q(a)
            g.set_target(b)
            g.set_params(params)
            gr, lv = g.grad(dx, dx, dtc, mode)
            params = opt.step(params, gr)
            hist["step"].append(lv)
            if log and it % 50 == 0:
                log(f"iter {it} loss {lv:.4e} ({time.time() - t0:.1f}s)")
            if (it + 1) % val_every == 0:
                validate(it + 1, params)
    if best[1] is not None:
        params = best[1]
    return params.astype(np.float32), hist


def one_step_loss(g, a, b, dx, dtc, params):
    """Mean squared one-step error of the unsplit step from a vs b, with the
    network (params) or with MC (params None)."""
    g.set_q(a)
    if params is None:
        g.bc2(MTHBC)
        g.step2(dx, dx, dtc, METHOD, [4, 4, 4])
    else:
        g.set_params(params)
        g.bc2(MTHBC)
        g.step2(dx, dx, dtc, METHOD, [6, 6, 6])
    q = g.get_q()[2:-2, 2:-2]
    return float(((q - b[2:-2, 2:-2]) ** 2).sum(axis=-1).mean())


def l1_error(q, ref, mbc=2):
    """Mean absolute difference per cell and field (interiors)."""
    return float(np.abs(q[mbc:-mbc, mbc:-mbc] - ref[mbc:-mbc, mbc:-mbc]).mean())


def run(n, seed, tend, limiter, params=None, arith=1):
    """Unsplit plain (limiter 4 = MC) or learned (6) run of the four-quadrant
    problem on n x n to time tend; returns (q, gpu_seconds, steps)."""
    q0, dx, dy = P.euler_quadrants(n, n, seed=seed)
    with LearnedGrid(SOLVER_EULER, n, n, [1.4, 1.0], arith=arith) as g:
        if params is not None:
            g.set_params(params)
        g.set_q(q0)
        t0 = time.perf_counter()
        out = g.claw2(dx, dy, 0.0, tend, [0.1 * dx, 1e99], [1.0, 0.9], 100000, METHOD,
                      [limiter] * 3, MTHBC)
        el = time.perf_counter() - t0
        return g.get_q(), el, int(out["nv"][1])


def evaluate(params, seeds=(200, 201), n=1024, tend=0.1, fine=4):
    """Config-5 comparison on held-out seeds: L1 error against the fine
    plain-MC run (n*fine)^2, measured at n^2 resolution, and GPU time of
    learned@n, MC@n and MC@2n.  Returns a list of dicts."""
    out = []
    for s in seeds:
        qf, tf, nf = run(n * fine, s, tend, 4)
        ref = coarsen(qf, fine)
        ql, tl, nl = run(n, s, tend, 6, params)
        qm, tm, nm = run(n, s, tend, 4)
        q2, t2, n2 = run(2 * n, s, tend, 4)
        out.append(dict(seed=s, tend=tend, l1_learned=l1_error(ql, ref), l1_mc=l1_error(qm, ref),
                        l1_mc2x=l1_error(coarsen(q2, 2), ref), t_learned=tl, t_mc=tm, t_mc2x=t2,
                        steps_learned=nl, steps_mc=nm, steps_mc2x=n2, t_fine=tf))
    return out


TRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_limiter_euler1024.npy")


def trained_params():
    """The network shipped with the package: trained by main() with its
    defaults (written to TRAINED with --save)."""
    p = np.load(TRAINED, allow_pickle=False)
    if p.shape != (NPARAMS,) or p.dtype != np.float32:
        raise ValueError(f"{TRAINED}: expected {NPARAMS} float32 parameters")
    return p


def main():
    """Train with the recipe of the shipped network (its defaults) and
    evaluate on held-out seeds: stage 1 regression onto superbee, stage 2
    the one-step squared error on pairs sampled over (0, 0.1] at Courant
    number 0.9 with the network kept by the multi-step error of a
    validation seed, stage 3 evolution strategies on the multi-step error of
    the training seeds over b2, W3 and b3, then (--es2-iters) over W2 as
    well, then (--es3-iters) the same over six training seeds."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train and evaluate the learned limiter")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[100, 101, 102])
    ap.add_argument("--val-seed", type=int, default=150)
    ap.add_argument("--eval-seeds", type=int, nargs="+", default=[200, 201])
    ap.add_argument("--nsnap", type=int, default=6)
    ap.add_argument("--pre-iters", type=int, default=300)
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--es-iters", type=int, default=30)
    ap.add_argument("--es2-iters", type=int, default=130)
    ap.add_argument("--es3-seeds", type=int, nargs="+", default=[100, 101, 102, 103, 104, 105])
    ap.add_argument("--es3-iters", type=int, default=52)
    ap.add_argument("--tend", type=float, default=0.1)
    ap.add_argument("--params", default=None, help="load parameters instead of training")
    ap.add_argument("--save", default=None)
    args = ap.parse_args()
    if args.params:
        params = np.load(args.params, allow_pickle=False).astype(np.float32)
        hist = None
    else:
        params, hist = train(args.n, tuple(args.train_seeds), args.nsnap, args.pre_iters,
                             args.iters, lr=2e-3, lr2=5e-4, courant=0.9, span=args.tend,
                             pretarget="superbee", val_seed=args.val_seed, val_every=10,
                             val_tend=args.tend, es_iters=args.es_iters, es_tend=args.tend,
                             log=lambda m: print(m, flush=True))
        if args.es2_iters > 0:
            params, h2 = train(args.n, tuple(args.train_seeds), 0, 0, 0, init=params,
                               es_iters=args.es2_iters, es_pairs=8, es_sigma=0.01, es_lr=0.005,
                               es_tend=args.tend, es_from=5 * H + H,
                               log=lambda m: print(m, flush=True))
            hist["es"] = hist["es"] + h2["es"]
        if args.es3_iters > 0:
            params, h3 = train(args.n, tuple(args.es3_seeds), 0, 0, 0, init=params,
                               es_iters=args.es3_iters, es_pairs=8, es_sigma=0.01, es_lr=0.005,
                               es_tend=args.tend, es_from=5 * H + H, seed=2,
                               log=lambda m: print(m, flush=True))
            hist["es"] = hist["es"] + h3["es"]
        if args.save:
            np.save(args.save, params)
    res = evaluate(params, tuple(args.eval_seeds), args.n, args.tend)
    rec = {"n": args.n, "results": res}
    if hist:
        rec["val"] = hist["val"]
        rec["val_mc"] = hist["val_mc"]
        rec["es"] = hist.get("es")
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
