"""Write tests/golden/claw2_hooks_<case>.npz: the reference claw2
(src/2d/claw2.f, compiled from /root/reference into oracle/_ref by
``make -C oracle ref``) run with the test user routines of
oracle/ref_shims.c (b4step2 and src2 logging their calls, src2 decaying
q(1)), on the cases of tests/hooks_cases.py.

    python tests/golden/make_golden_hooks.py
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
from tests import hooks_cases as H  # noqa: E402


def main():
    O.build(ref=True)
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


if __name__ == "__main__":
    main()
