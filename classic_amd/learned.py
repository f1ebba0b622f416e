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


def fine_snapshots(n_coarse, seed, nsnap, factor=4, cfl=0.8, arith=1):
    """Run the plain method on the fine grid (factor * n_coarse)^2 with a
    fixed dt at Courant number `cfl` and return coarsened snapshots
    [(q_t, q_t+dtc)] separated by `factor` fine steps (one coarse step of
    dtc = factor * dt_fine), plus the coarse dx and dtc."""
    nf = n_coarse * factor
    q0, dxf, dyf = P.euler_quadrants(nf, nf, seed=seed)
    # max wave speed of the initial data
    rho, u, v = q0[2:-2, 2:-2, 0], q0[2:-2, 2:-2, 1] / q0[2:-2, 2:-2, 0], q0[2:-2, 2:-2, 2] / q0[2:-2, 2:-2, 0]
    p = 0.4 * (q0[2:-2, 2:-2, 3] - 0.5 * rho * (u * u + v * v))
    smax = (np.sqrt(u * u + v * v) + np.sqrt(1.4 * p / rho)).max()
    dtf = cfl * dxf / (smax * 1.6)
    out = []
    with Grid2D(SOLVER_EULER, nf, nf, [1.4, 1.0], arith=arith) as g:
        g.set_q(q0)
        # let the waves develop before sampling
        g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, 8 * factor)
        for _ in range(nsnap):
            a = coarsen(g.get_q(), factor)
            g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, factor)
            b = coarsen(g.get_q(), factor)
            out.append((a, b))
            g.time_steps(dxf, dyf, dtf, METHOD, [4, 4, 4], MTHBC, 2 * factor)
    return out, dxf * factor, dtf * factor


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
          seed=0, log=None):
    """Train the network: (1) regression onto MC (the network starts as the
    plain method), (2) the one-step loss of the unsplit step against
    4x-coarsened fine solutions.  Returns (params, history)."""
    data = []
    for s in seeds:
        snaps, dx, dtc = fine_snapshots(n_coarse, s, nsnap)
        data += [(a, b, dx, dtc) for a, b in snaps]
    params = init_params(seed).astype(np.float64)
    hist = {"pre": [], "step": [], "step_mc": []}
    t0 = time.time()
    with LearnedGrid(SOLVER_EULER, n_coarse, n_coarse, [1.4, 1.0], arith=1) as g:
        # stage 1: MC regression on the features of all snapshots
        opt = Adam(NPARAMS, lr=lr)
        for it in range(pre_iters):
            a, b, dx, dtc = data[it % len(data)]
            g.set_q(a)
            g.bc2(MTHBC)
            g.features(dx, dx, dtc)
            nlive = max(1, int((g.read(0)[:, 3] > 0).sum()))
            g.set_params(params)
            gr, loss = g.grad(dx, dx, dtc, 0)
            params = opt.step(params, gr / nlive)
            hist["pre"].append(loss / nlive)
        # baseline one-step loss of MC on every snapshot
        for a, b, dx, dtc in data:
            hist["step_mc"].append(one_step_loss(g, a, b, dx, dtc, None))
        # stage 2: one-step loss
        opt = Adam(NPARAMS, lr=lr * 0.5)
        for it in range(iters):
            a, b, dx, dtc = data[it % len(data)]
            g.set_q(a)
            g.set_target(b)
            g.set_params(params)
            gr, loss = g.grad(dx, dx, dtc, 1)
            params = opt.step(params, gr)
            hist["step"].append(loss)
            if log and it % 50 == 0:
                log(f"iter {it} loss {loss:.4e} ({time.time() - t0:.1f}s)")
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


def main():
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description="train and evaluate the learned limiter")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[100, 101, 102, 103])
    ap.add_argument("--eval-seeds", type=int, nargs="+", default=[200, 201])
    ap.add_argument("--nsnap", type=int, default=4)
    ap.add_argument("--pre-iters", type=int, default=300)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--tend", type=float, default=0.1)
    ap.add_argument("--params", default=None, help="load parameters instead of training")
    ap.add_argument("--save", default=None)
    args = ap.parse_args()
    if args.params:
        params = np.load(args.params).astype(np.float32)
        hist = None
    else:
        params, hist = train(args.n, tuple(args.train_seeds), args.nsnap, args.pre_iters,
                             args.iters, log=lambda m: print(m, flush=True))
        if args.save:
            np.save(args.save, params)
    res = evaluate(params, tuple(args.eval_seeds), args.n, args.tend)
    rec = {"n": args.n, "results": res}
    if hist:
        rec["train_loss_first_last"] = [hist["step"][0], hist["step"][-1]]
        rec["train_mc_loss_mean"] = float(np.mean(hist["step_mc"]))
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
