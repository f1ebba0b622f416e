"""Write tests/golden/claw{1,2,3}_hooks_<case>.npz: the reference claw1 /
claw2 / claw3 (src/1d/claw1.f, src/2d/claw2.f, src/3d/claw3.f, compiled from
/root/reference into oracle/_ref by ``make -C oracle ref``) run with the test
user routines of oracle/ref_shims1.c, ref_shims.c and ref_shims3.c (b4step
and src logging their calls, src decaying q(1)), on the cases of
tests/hooks_cases.py.

    python tests/golden/make_golden_hooks.py [1] [2] [3]   (default: all)
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import oracle1 as O1  # noqa: E402
from oracle import oracle3 as O3  # noqa: E402
from tests import hooks_cases as H  # noqa: E402


def main2():
    R = O.ref()
    for name, c in H.CASES.items():
        q, dx, dy = H.initial(name)
        q = np.ascontiguousarray(q).copy()
        if c["solver"] == "euler":
            p = O.Problem(4, 3, c["mx"], c["my"], O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                          method=c["method"], mthlim=[4, 4, 4], mthbc=c["mthbc"])
        else:
            p = O.Problem(3, 2, c["mx"], c["my"], O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                          method=c["method"], mthlim=[4, 4], mthbc=c["mthbc"])
        R.ref_set_solver(p.solver_id, C.byref(p.params))
        log = np.zeros((4096, 3))
        R.ref_hooks_set(C.c_double(c["k"]), log.ctypes.data_as(C.POINTER(C.c_double)), 4096)
        te = C.c_double(c["tend"])
        dtv = np.zeros(5)
        dtv[:2] = c["dtv"]
        cflv = np.array([1.0, 0.9, 0.0, 0.0])
        nv = np.array([500, 0], dtype=np.int32)
        info = R.ref_claw2_hooks(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, O._dp(q), None,
                                 C.c_double(0.0), C.c_double(0.0), C.c_double(dx), C.c_double(dy),
                                 C.c_double(0.0), C.byref(te), O._dp(dtv), O._dp(cflv), O._ip(nv),
                                 O._ip(p.method), O._ip(p.mthlim), O._ip(p.mthbc), p.use_fwave)
        n = R.ref_hooks_count()
        q0, _, _ = H.initial(name)
        np.savez_compressed(os.path.join(HERE, f"claw2_hooks_{name}.npz"), q0=q0, q=q,
                            tend=te.value, dtv=dtv, cflv=cflv, nv=nv, info=info, log=log[:n])
        kinds = log[:n, 0]
        print(name, "info", info, "steps", nv[1], "b4step2 calls", int((kinds == 1).sum()),
              "src2 calls", int((kinds == 2).sum()))


def _summary(name, info, nv, log):
    kinds = log[:, 0]
    print(name, "info", info, "steps", nv[1], "b4step calls", int((kinds == 1).sum()),
          "src calls", int((kinds == 2).sum()))


def main1():
    R = O1.ref()
    for name, c in H.CASES1.items():
        q, aux, dx = H.initial1(name)
        q = np.ascontiguousarray(q).copy()
        vc = c["solver"] == "vc_acoustics"
        p = O1.Problem1(c["mx"], O1.SOLVER_VC_ACOUSTICS if vc else O1.SOLVER_ACOUSTICS,
                        O1.params1(1.3, 2.9), maux=c["maux"], method=c["method"],
                        mthlim=c["mthlim"], mthbc=c["mthbc"])
        O1._setp(p, "reference")
        log = np.zeros((4096, 3))
        R.ref1_hooks_set(C.c_double(c["k"]), log.ctypes.data_as(C.POINTER(C.c_double)), 4096)
        te = C.c_double(c["tend"])
        dtv = np.zeros(5)
        dtv[:2] = c["dtv"]
        cflv = np.array([1.0, 0.9, 0.0, 0.0])
        nv = np.array([500, 0], dtype=np.int32)
        info = R.ref1_claw1_hooks(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, O1._dp(q),
                                  O1._dp(aux), C.c_double(0.0), C.c_double(dx), C.c_double(0.0),
                                  C.byref(te), O1._dp(dtv), O1._dp(cflv), O1._ip(nv),
                                  O1._ip(p.method), O1._ip(p.mthlim), O1._ip(p.mthbc),
                                  p.use_fwave)
        n = R.ref1_hooks_count()
        q0, aux0, _ = H.initial1(name)
        extra = {} if aux0 is None else dict(aux=aux0)
        np.savez_compressed(os.path.join(HERE, f"claw1_hooks_{name}.npz"), q0=q0, q=q,
                            tend=te.value, dtv=dtv, cflv=cflv, nv=nv, info=info, log=log[:n],
                            **extra)
        _summary(name, info, nv, log[:n])


def main3():
    R = O3.ref()
    for name, c in H.CASES3.items():
        q, aux = H.initial3(name)
        q = np.ascontiguousarray(q).copy()
        p = O3.Problem3(c["mx"], c["my"], c["mz"], method=c["method"], mthlim=c["mthlim"],
                        mthbc=c["mthbc"])
        log = np.zeros((4096, 3))
        R.ref3_hooks_set(C.c_double(c["k"]), log.ctypes.data_as(C.POINTER(C.c_double)), 4096)
        te = C.c_double(c["tend"])
        dtv = np.zeros(5)
        dtv[:2] = c["dtv"]
        cflv = np.array([1.0, 0.9, 0.0, 0.0])
        nv = np.array([500, 0], dtype=np.int32)
        info = R.ref3_claw3_hooks(p.meqn, p.mwaves, p.maux, p.mbc, p.mx, p.my, p.mz, O3._dp(q),
                                  O3._dp(aux), C.byref(p.geom), C.c_double(0.0), C.byref(te),
                                  O3._dp(dtv), O3._dp(cflv), O3._ip(nv), O3._ip(p.method),
                                  O3._ip(p.mthlim), O3._ip(p.mthbc), p.use_fwave)
        n = R.ref3_hooks_count()
        q0, _ = H.initial3(name)
        np.savez_compressed(os.path.join(HERE, f"claw3_hooks_{name}.npz"), q0=q0, q=q,
                            tend=te.value, dtv=dtv, cflv=cflv, nv=nv, info=info, log=log[:n])
        _summary(name, info, nv, log[:n])


def main(dims):
    O.build(ref=True)
    for d, run in (("1", main1), ("2", main2), ("3", main3)):
        if d in dims:
            run()


if __name__ == "__main__":
    main(sys.argv[1:] or ["1", "2", "3"])
