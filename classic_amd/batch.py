"""Batched 2D problems on the device (``clawgpu_batch_*``).

``Batch2D`` holds many independent grids of different sizes and solvers and
advances each with its own claw2 loop (src/2d/claw2.f:494-784): the same
dt sequence, Courant checks, retakes and results as ``Grid2D.claw2`` on that
grid alone, bit for bit in exact arithmetic.  One batched iteration makes one
step attempt on every unfinished problem in a single persistent kernel;
iterations are replayed from a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SOLVER_ACOUSTICS, SOLVER_EULER, check
from .gpu import _dp, _ip, _ivec

_SOLVER_IDS = {"acoustics": SOLVER_ACOUSTICS, "euler": SOLVER_EULER,
               SOLVER_ACOUSTICS: SOLVER_ACOUSTICS, SOLVER_EULER: SOLVER_EULER}
_MEQN = {SOLVER_ACOUSTICS: 3, SOLVER_EULER: 4}


class Batch2D:
    """Device-resident batch of 2D problems.

    specs: list of dicts with keys solver ("acoustics"/"euler" or the solver
    id), mx, my, params (rho, bulk) / (gamma, efix), mthbc (4 ints), dx, dy.
    """

    def __init__(self, specs, mbc=2, arith=0):
        self.lib = _lib.load()
        self.n = len(specs)
        self.mbc = mbc
        self.solver = np.array([_SOLVER_IDS[s["solver"]] for s in specs], dtype=np.int32)
        self.mx = np.array([s["mx"] for s in specs], dtype=np.int32)
        self.my = np.array([s["my"] for s in specs], dtype=np.int32)
        params = np.ascontiguousarray(np.array([list(s["params"])[:2] for s in specs],
                                               dtype=np.float64).reshape(-1))
        mthbc = _ivec(np.array([list(s["mthbc"]) for s in specs]).reshape(-1), 4 * self.n)
        self.dx = np.ascontiguousarray([float(s["dx"]) for s in specs], dtype=np.float64)
        self.dy = np.ascontiguousarray([float(s["dy"]) for s in specs], dtype=np.float64)
        h = C.c_void_p()
        check(self.lib.clawgpu_batch_create(C.byref(h), self.n, mbc, _ip(self.solver), _ip(self.mx),
                                            _ip(self.my), _dp(params), _ip(mthbc), _dp(self.dx),
                                            _dp(self.dy)))
        self.h = h
        if arith:
            self.set_arith(arith)

    def close(self):
        if getattr(self, "h", None):
            self.lib.clawgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def shape(self, k):
        return (int(self.my[k]) + 2 * self.mbc, int(self.mx[k]) + 2 * self.mbc,
                _MEQN[int(self.solver[k])])

    def cells(self, k=None):
        if k is None:
            return int(np.sum(self.mx.astype(np.int64) * self.my))
        return int(self.mx[k]) * int(self.my[k])

    def set_q(self, k, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.shape(k):
            raise ValueError(f"q has shape {q.shape}, expected {self.shape(k)}")
        check(self.lib.clawgpu_batch_set_q(self.h, int(k), _dp(q)))

    def get_q(self, k):
        q = np.empty(self.shape(k))
        check(self.lib.clawgpu_batch_get_q(self.h, int(k), _dp(q)))
        return q

    def set_arith(self, mode):
        check(self.lib.clawgpu_batch_set_arith(self.h, int(mode)))

    def start(self, tstart, tend, dtv, cflv, nmax, method, mthlim, use_fwave=False):
        """Begin the claw2 runs: tend is a scalar or one value per problem."""
        te = np.ascontiguousarray(np.broadcast_to(np.asarray(tend, dtype=np.float64), (self.n,)))
        dtv_a = np.zeros(5)
        dtv_a[: len(dtv)] = dtv
        cflv_a = np.zeros(4)
        cflv_a[: len(cflv)] = cflv
        lim = np.zeros(3, dtype=np.int32)
        ml = np.asarray(mthlim, dtype=np.int32).reshape(-1)
        lim[: ml.size] = ml[:3]
        if ml.size < 3:
            lim[ml.size:] = ml[-1]
        check(self.lib.clawgpu_batch_start(self.h, float(tstart), _dp(te), _dp(dtv_a), _dp(cflv_a),
                                           int(nmax), _ip(_ivec(method, 7)), _ip(lim),
                                           int(bool(use_fwave))))
        self._nmax = int(nmax)

    def advance(self, max_iters=None, graph=True):
        """Run batched iterations until every problem is done or max_iters
        (default: three per allowed step, enough for a retake after every
        step); returns the number of unfinished problems."""
        if max_iters is None:
            max_iters = min(3 * getattr(self, "_nmax", 0) + 16, (1 << 31) - 1)
        left = C.c_int(0)
        check(self.lib.clawgpu_batch_advance(self.h, int(max_iters), int(bool(graph)),
                                             C.byref(left)))
        return left.value

    def result(self, k):
        """claw2's outputs of problem k (as Grid2D.claw2 returns them)."""
        t = C.c_double(0.0)
        dtv = np.zeros(5)
        cflv = np.zeros(4)
        nv = np.zeros(2, dtype=np.int32)
        info = C.c_int(0)
        check(self.lib.clawgpu_batch_result(self.h, int(k), C.byref(t), _dp(dtv), _dp(cflv),
                                            _ip(nv), C.byref(info)))
        return dict(info=info.value, tend=t.value, dtv=dtv, cflv=cflv, nv=nv)

    def set_schedule(self, seglen, nblocks=0, largest_first=True, w3=False):
        """Rows per tile, blocks of each persistent step kernel (0: two per
        CU), tile order, and the acoustics kernel's three-waves budget."""
        s = np.array([seglen, nblocks, int(bool(largest_first)), int(bool(w3))], dtype=np.int32)
        check(self.lib.clawgpu_batch_set_schedule(self.h, _ip(s)))

    def schedule(self):
        s = np.zeros(4, dtype=np.int32)
        check(self.lib.clawgpu_batch_get_schedule(self.h, _ip(s)))
        return dict(seglen=int(s[0]), nblocks=int(s[1]), largest_first=bool(s[2]), w3=bool(s[3]))

    def tune(self, reps=3, max_cand=128):
        """Autotune the step kernels' schedule on state-preserving
        iterations; returns the measured candidates as
        (seglen, nblocks, largest_first, w3, ms), the fastest of which is
        now in use."""
        nc = C.c_int(0)
        cand = np.zeros(4 * max_cand, dtype=np.int32)
        ms = np.zeros(max_cand, dtype=np.float32)
        check(self.lib.clawgpu_batch_tune(self.h, int(reps), C.byref(nc), _ip(cand),
                                          ms.ctypes.data_as(C.POINTER(C.c_float)), max_cand))
        n = min(nc.value, max_cand)
        return [(int(cand[4 * i]), int(cand[4 * i + 1]), bool(cand[4 * i + 2]),
                 bool(cand[4 * i + 3]), float(ms[i])) for i in range(n)]

    def time_iterations(self, iters):
        """iters batched iterations without the graph, timed with HIP
        events; returns (total_ms, step_kernel_ms)."""
        tot, ker = C.c_float(0.0), C.c_float(0.0)
        check(self.lib.clawgpu_batch_time(self.h, int(iters), C.byref(tot), C.byref(ker)))
        return tot.value, ker.value


def from_specs(specs, mbc=2, arith=0):
    """Batch2D of problems.batch_problem_specs entries with their initial
    data loaded; returns (batch, list of initial q)."""
    from . import problems as P

    full, qs = [], []
    for s in specs:
        q, dx, dy = P.batch_problem_q(s, mbc=mbc)
        full.append(dict(s, dx=dx, dy=dy))
        qs.append(q)
    b = Batch2D(full, mbc=mbc, arith=arith)
    for k, q in enumerate(qs):
        b.set_q(k, q)
    return b, qs
