"""Write the reference driver outputs of the driver cases under
tests/golden/claw2ez_<case>/ (claw.data, fort.q/t/a files and fort.info,
gzip-compressed).

Runs the reference drivers src/2d/claw2ez.f, src/1d/claw1ez.f and
src/3d/claw3ez.f (with the claw loops, output routines and the rest of the
reference sources, compiled from /root/reference into oracle/_ref by
``make -C oracle ref``) on claw.data files written by
classic_amd.rundata.ClawRunData, with qinit returning the case's initial
data.  The GPU driver test compares its own files with these byte for byte.

    python tests/golden/make_golden_ez.py
"""
from __future__ import annotations

import gzip
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import oracle1 as O1  # noqa: E402
from oracle import oracle3 as O3  # noqa: E402
from tests import ez_cases as E  # noqa: E402


def _run(name):
    """Run the reference driver of a case in a temporary directory; yields it."""
    if name in E.CASES:
        rd, q0, solver, params = E.CASES[name]()
        sid = O.SOLVER_BY_NAME[solver]
        sp = O.solver_params(*params) if solver == "acoustics" else O.solver_params(
            gamma=params[0], efix=bool(params[1]))
        return rd, lambda d: O.ref_claw2ez(d, sid, sp, q0)
    if name in E.CASES1:
        rd, q0, _, params = E.CASES1[name]()
        return rd, lambda d: O1.ref_claw1ez(d, O1.SOLVER_ACOUSTICS, O1.params1(*params), q0)
    rd, (q0, aux0), _, _ = E.CASES3[name]()
    return rd, lambda d: O3.ref_claw3ez(d, q0, aux0)


def main():
    O.build(ref=True)
    for name in E.GOLDEN + E.GOLDEN1 + E.GOLDEN3:
        rd, run = _run(name)
        out = os.path.join(HERE, f"claw2ez_{name}")
        shutil.rmtree(out, ignore_errors=True)
        os.makedirs(out)
        with tempfile.TemporaryDirectory() as tmp:
            rd.write(tmp)
            run(tmp)
            for f in sorted(os.listdir(tmp)):
                with open(os.path.join(tmp, f), "rb") as src, \
                        open(os.path.join(out, f + ".gz"), "wb") as raw, \
                        gzip.GzipFile(fileobj=raw, mode="wb", mtime=0, filename="") as dst:
                    dst.write(src.read())
        print(name, sorted(os.listdir(out)))


if __name__ == "__main__":
    main()
