"""Write the 1D golden vectors under tests/golden/ from the REFERENCE build.

Runs the reference Fortran (clawpack/classic src/1d: claw1, step1,
inlinelimiter, bc1), compiled from /root/reference into oracle/_ref by
``make -C oracle ref`` and linked with this project's rp1 solvers, on the two
acoustics examples and on single steps, and copies the examples' regression
frame sums (examples/acoustics_1d_*/regression_data/frame000{1,2}.txt).

    python tests/golden/make_golden1.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from classic_amd import problems as P  # noqa: E402
from oracle import oracle1 as O1  # noqa: E402

EX = "/root/reference/examples"


def regression(name):
    return np.array([np.loadtxt(os.path.join(EX, name, "regression_data", f"frame000{k}.txt"))
                     for k in (1, 2)])


def main():
    if not O1.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {}
    q, dx = P.acoustics1d_gaussian(400)
    p = O1.Problem1(400, O1.SOLVER_ACOUSTICS, O1.params1(1.0, 4.0), mthlim=(4, 4))
    fr = O1.frames(p, q, dx, 0.0, 1.0, 2, 0.1, nmax=500, backend="ref")
    out["acoustics1d_example1"] = dict(t=np.array([t for t, _ in fr]),
                                       q=np.array([qq for _, qq in fr]),
                                       regression_sums=regression("acoustics_1d_example1"))
    q, aux, dx = P.acoustics1d_heterogeneous(500)
    p = O1.Problem1(500, O1.SOLVER_VC_ACOUSTICS, maux=2, mthlim=(4, 4))
    fr = O1.frames(p, q, dx, 0.0, 5.0, 2, 1.0, dt_max=1e9, aux=aux, backend="ref")
    out["acoustics1d_heterogeneous"] = dict(t=np.array([t for t, _ in fr]),
                                            q=np.array([qq for _, qq in fr]),
                                            regression_sums=regression("acoustics_1d_heterogeneous"))
    rng = np.random.default_rng(11)
    cases = [(O1.SOLVER_ACOUSTICS, 0, (1, 2, 0, 0, 0, 0, 0), (1, 2), 0),
             (O1.SOLVER_ACOUSTICS, 0, (1, 2, 0, 0, 0, 0, 0), (3, 4), 1),
             (O1.SOLVER_VC_ACOUSTICS, 2, (1, 2, 0, 0, 0, 0, 0), (4, 2), 0),
             (O1.SOLVER_VC_ACOUSTICS, 3, (1, 2, 0, 0, 0, 3, 0), (2, 3), 0),
             (O1.SOLVER_ADVECTION, 0, (1, 2, 0, 0, 0, 0, 0), (4,), 0),
             (O1.SOLVER_ADVECTION, 0, (1, 1, 0, 0, 0, 0, 0), (0,), 0)]
    for n, (sid, maux, method, lim, fw) in enumerate(cases):
        mw = 1 if sid == O1.SOLVER_ADVECTION else 2
        p = O1.Problem1(57, sid, O1.params1(1.3, 2.9, u=-0.7), meqn=mw, mwaves=mw, maux=maux,
                        method=method, mthlim=lim, use_fwave=bool(fw))
        q = rng.standard_normal((61, mw))
        aux = None
        if maux:
            aux = np.stack([rng.uniform(0.5, 3, 61), rng.uniform(0.5, 2, 61)] +
                           ([rng.uniform(0.5, 1.5, 61)] if maux > 2 else []), -1)
        qo, cfl = O1.step1(p, q, 2.0 / 57, 0.01, aux=aux, backend="ref")
        out[f"step1_case{n}"] = dict(q_in=q, aux=aux if aux is not None else np.zeros(0),
                                     q_out=qo, cfl=cfl, solver=sid, method=np.array(p.method),
                                     mthlim=np.array(lim), use_fwave=fw)
    for k, v in out.items():
        np.savez_compressed(os.path.join(HERE, k + ".npz"), **v)
        print("wrote", k)


if __name__ == "__main__":
    main()
