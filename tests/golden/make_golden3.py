"""Write the 3D golden vectors under tests/golden/ from the REFERENCE build.

Runs the reference Fortran (clawpack/classic src/3d: claw3, step3, flux3,
limiter/philim, bc3, dimsp3, step3ds, and the heterogeneous-acoustics
example's bc3.f), compiled from /root/reference into oracle/_ref by
``make -C oracle ref`` and linked with this project's variable-coefficient
acoustics solvers, on small problems, and stores inputs and outputs as .npz
fixtures.  It also copies the example's own regression frame sums
(examples/acoustics_3d_heterogeneous/regression_data/frame000{1,2}.txt).

    python tests/golden/make_golden3.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from classic_amd import problems as P  # noqa: E402
from oracle import oracle3 as O3  # noqa: E402

REGRESSION = "/root/reference/examples/acoustics_3d_heterogeneous/regression_data"

METHODS = [((1, 2, 22), (3, 3), 0, 0), ((1, 2, 21), (1, 2), 0, 0), ((1, 2, 20), (4, 4), 0, 0),
           ((1, 1, 11), (0, 0), 0, 0), ((1, 1, 10), (0, 0), 0, 0), ((1, 2, 0), (2, 3), 0, 0),
           ((1, 2, 22), (4, 1), 3, 0), ((1, 2, 22), (3, 3), 0, 1), ((1, 2, -1), (3, 3), 0, 0),
           ((1, 2, -1), (4, 4), 3, 0)]


def random_state(p, rng, maux):
    q = rng.standard_normal(p.shape())
    aux = np.empty(p.shape(maux))
    aux[..., 0] = rng.uniform(0.5, 3.0, aux.shape[:3])
    aux[..., 1] = rng.uniform(0.5, 2.0, aux.shape[:3])
    if maux > 2:
        aux[..., 2] = rng.uniform(0.5, 1.5, aux.shape[:3])
    return q, aux


def main():
    if not O3.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {}
    # 1) the example's regression configuration: 20^3, frames at t = 0.5, 1
    p = O3.Problem3(20, 20, 20, mthbc=(1, 1, 1, 1, 1, 0), mthlim=(3, 3))
    q, aux = P.acoustics3d_heterogeneous(20)
    fr = O3.frames(p, q, aux, 0.0, 1.0, 2, 0.01, backend="ref")
    reg = np.array([np.loadtxt(os.path.join(REGRESSION, f"frame000{k}.txt")) for k in (1, 2)])
    out["hetero3d_frames_20"] = dict(
        t=np.array([t for t, _ in fr]),
        sums=np.array([[qq[2:-2, 2:-2, 2:-2, m].sum() for m in range(4)] for _, qq in fr]),
        q_frame1_sub=fr[0][1][2:-2:3, 2:-2:3, 2:-2:3].copy(),
        regression_sums=reg)

    # 2) single steps, every method family, random states and materials
    rng = np.random.default_rng(1)
    for n, (meth, lim, capa, fw) in enumerate(METHODS):
        maux = 3 if capa else 2
        p = O3.Problem3(6, 5, 4, maux=maux, mthlim=lim,
                        method=(meth[0], meth[1], meth[2], 0, 0, capa, 0), use_fwave=bool(fw))
        q, aux = random_state(p, rng, maux)
        if meth[2] >= 0:
            qo, cfl = O3.step3(p, q, aux, 0.02, backend="ref")
        else:
            qo, cfl = O3.dimsp3(p, q, aux, 0.02, backend="ref")
        out[f"step3_case{n}"] = dict(q_in=q, aux=aux, q_out=qo, cfl=cfl,
                                     method=np.array(p.method), mthlim=np.array(lim),
                                     use_fwave=fw)

    # 3) bc3 with every code and the piston
    for n, mthbc in enumerate([(1, 1, 1, 1, 1, 1), (2, 2, 2, 2, 2, 2), (3, 3, 3, 3, 3, 3),
                               (1, 3, 2, 2, 3, 0)]):
        p = O3.Problem3(5, 4, 6, mthbc=mthbc, lower=(-0.2, -0.1, -1.0))
        q = rng.standard_normal(p.shape())
        out[f"bc3_case{n}"] = dict(q_in=q, mthbc=np.array(mthbc), t=0.03,
                                   q_out=O3.bc3(p, q, 0.03, backend="ref"))

    # 4) a claw3 run with retakes: 12x10x8 pulse in a layered medium, walls
    p = O3.Problem3(12, 10, 8, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2))
    qc, _, _ = P.acoustics3d_pulse(12, seed=2, width=0.4)
    q = qc[: p.shape()[0], : p.shape()[1], : p.shape()[2]].copy()
    aux = P.layered_medium_aux(12, 10, 8, below=(1.0, 1.0), above=(6.0, 2.0))
    r = O3.claw3(p, q, aux, 0.0, 0.5, [0.2, 1e9], [1.0, 0.9], 500, backend="ref")
    out["claw3_pulse_12x10x8"] = dict(q_in=q, aux=aux, q_out=r.q, info=r.info, tend=r.tend,
                                      dtv=r.dtv, cflv=r.cflv, nv=r.nv)

    for k, v in out.items():
        np.savez_compressed(os.path.join(HERE, k + ".npz"), **v)
        print("wrote", k)


if __name__ == "__main__":
    main()
