"""Write the golden vectors under tests/golden/ from the REFERENCE build.

Runs the reference Fortran (clawpack/classic src/2d: claw2, step2, flux2,
inlinelimiter, bc2, dimsp2, step2ds), compiled from /root/reference into
oracle/_ref by ``make -C oracle ref`` and linked with this project's Riemann
solvers, on small problems, and stores inputs and outputs as .npz fixtures.
The fixtures let the CPU tests pin the oracle on machines where the
reference tree or flang is absent.

    python tests/golden/make_golden.py [name ...]   (all fixtures when no name is given)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from classic_amd import problems as P  # noqa: E402
from oracle import oracle as O  # noqa: E402

GAMMA = 1.4
REF_EX = "/root/reference/examples"


def acoustics_problem(mx, my, method=(1, 2, 2, 0, 0, 0, 0), lim=(4, 4), mthbc=(3, 3, 3, 3),
                      rho=1.0, bulk=4.0):
    return O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(rho, bulk), method=method,
                     mthlim=list(lim), mthbc=mthbc)


def euler_problem(mx, my, method=(1, 2, 2, 0, 0, 0, 0), lim=(4, 4, 4), mthbc=(1, 1, 1, 1),
                  efix=True):
    return O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=GAMMA, efix=efix),
                     method=method, mthlim=list(lim), mthbc=mthbc)


def claw_record(p, q, dx, dy, tend, dtv, cflv, nmax):
    r = O.claw2(p, q, dx, dy, 0.0, tend, dtv, cflv, nmax, backend="ref")
    return dict(q_out=r.q[2:-2, 2:-2], info=r.info, tend=r.tend, dtv=r.dtv, cflv=r.cflv,
                nv=r.nv)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {}

    # 1) config-1 geometry at 40x40: acoustics pulse, walls, MC, to t = 0.3
    q, dx, dy = P.acoustics_pulse(40, 40)
    p = acoustics_problem(40, 40)
    out["acoustics_pulse_40"] = dict(q_in=q, dx=dx, dy=dy, tend_in=0.3,
                                     **claw_record(p, q, dx, dy, 0.3, [1e-3, 1e99], [1.0, 0.9],
                                                   50000))

    # 2) config-1 itself (200x200 to t = 0.6): sums and a subsample only
    q, dx, dy = P.acoustics_pulse(200, 200)
    p = acoustics_problem(200, 200)
    r = claw_record(p, q, dx, dy, 0.6, [1e-3, 1e99], [1.0, 0.9], 50000)
    qo = r.pop("q_out")
    out["config1_acoustics_200"] = dict(sums=qo.sum(axis=(0, 1)), l1=np.abs(qo).sum(axis=(0, 1)),
                                        sub=qo[::10, ::10].copy(), **r)

    # 3) Euler quadrants 32x24, 20 steps, MC, extrapolation
    q, dx, dy = P.euler_quadrants(32, 24, seed=3)
    p = euler_problem(32, 24)
    out["euler_quadrants_32x24"] = dict(q_in=q, dx=dx, dy=dy,
                                        **claw_record(p, q, dx, dy, 1e9, [1e-4, 1e99], [1.0, 0.9],
                                                      20))

    # 4) single unsplit steps on a ghost-filled Euler state: every method and
    #    limiter, walls on x, periodic on y
    q, dx, dy = P.euler_quadrants(17, 13, seed=8)
    rec = {"q_in": None, "dx": dx, "dy": dy, "dt": 0.25 * dx}
    for meth in [(1, 2, 2), (1, 2, 1), (1, 2, 0), (1, 1, 1), (1, 1, 0)]:
        for lim in [0, 1, 2, 3, 4, 5]:
            p = euler_problem(17, 13, method=meth + (0, 0, 0, 0), lim=(lim,) * 3,
                              mthbc=(3, 3, 2, 2))
            qb = O.bc2(p, q, backend="ref")
            rec["q_in"] = qb
            qn, cfl = O.step2(p, qb, dx, dy, 0.25 * dx, backend="ref")
            key = f"m{meth[1]}{meth[2]}_l{lim}"
            rec["q_" + key] = qn[2:-2, 2:-2]
            rec["cfl_" + key] = cfl
    out["euler_step2_17x13"] = rec

    # 5) dimensional splitting (Godunov and Strang), Euler 24x20, 10 steps
    q, dx, dy = P.euler_quadrants(24, 20, seed=5)
    for split in (-1, -2):
        p = euler_problem(24, 20, method=(1, 2, split, 0, 0, 0, 0))
        out[f"euler_dimsplit{-split}_24x20"] = dict(
            q_in=q, dx=dx, dy=dy,
            **claw_record(p, q, dx, dy, 1e9, [1e-4, 1e99], [1.0, 0.9], 10))

    # 6) bc2 on every boundary-condition combination
    q = np.random.default_rng(0).standard_normal((13 + 4, 11 + 4, 3))
    bc = {"q_in": q}
    for mthbc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2), (3, 1, 3, 1)]:
        p = acoustics_problem(11, 13, mthbc=mthbc)
        bc["q_" + "".join(map(str, mthbc))] = O.bc2(p, q, backend="ref")
    out["bc2_11x13"] = bc

    # 7) the annulus example (variable-coefficient advection with capacity):
    #    frames at t = 0.25, 0.5 as its regression test runs it, and the
    #    regression sums its directory holds
    aux, dxc, dyc = O.annulus_aux(40, 120)
    q = P.annulus_hump(40, 120)
    p = O.Problem(1, 1, 40, 120, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3,
                  method=(1, 2, 2, 0, 0, 3, 3), mthlim=[3], mthbc=(1, 1, 2, 2))
    frames, dtv, t = [], [0.1, 1e9], 0.0
    for _ in range(2):
        r = O.claw2(p, q, dxc, dyc, t, t + 0.25, dtv, [1.0, 0.9], 50000, aux=aux, backend="ref")
        q, dtv, t = r.q, [r.dtv[4], 1e9], r.tend
        frames.append(q[2:-2, 2:-2].copy())
    reg = np.array([np.loadtxt(os.path.join(REF_EX, "advection_2d_annulus", "regression_data",
                                            f"frame000{k}.txt")) for k in (1, 2)])
    out["annulus_40x120"] = dict(q=np.array(frames), t=np.array([0.25, 0.5]),
                                 regression_sums=reg)

    # 8) the annulus example taken by dimensional splitting (method(3) = -1
    #    Godunov, -2 Strang) with its capacity function, to t = 0.25
    split = {}
    for m3 in (-1, -2):
        p = O.Problem(1, 1, 40, 120, O.SOLVER_VC_ADVECTION, O.solver_params(), maux=3,
                      method=(1, 2, m3, 0, 0, 3, 3), mthlim=[3], mthbc=(1, 1, 2, 2))
        r = O.claw2(p, P.annulus_hump(40, 120), dxc, dyc, 0.0, 0.25, [0.1, 1e9], [1.0, 0.9],
                    50000, aux=aux, backend="ref")
        key = f"split{-m3}"
        split.update({f"q_{key}": r.q[2:-2, 2:-2], f"dtv_{key}": r.dtv, f"nv_{key}": r.nv,
                      f"info_{key}": r.info, f"tend_{key}": r.tend})
    out["annulus_split_40x120"] = split

    want = set(sys.argv[1:])  # names to write (all when none are given)
    for name, rec in out.items():
        if want and name not in want:
            continue
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            **{k: np.asarray(v) for k, v in rec.items()})
        print("wrote", name)


if __name__ == "__main__":
    main()
