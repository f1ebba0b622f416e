"""Write the reference claw2ez outputs of the driver cases under
tests/golden/claw2ez_<case>/ (claw.data, fort.q/t files and fort.info,
gzip-compressed).

Runs the reference driver src/2d/claw2ez.f (with claw2, out2 and the rest
of src/2d, compiled from /root/reference into oracle/_ref by
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
from tests import ez_cases as E  # noqa: E402


def main():
    O.build(ref=True)
    for name in E.GOLDEN:
        rd, q0, solver, params = E.CASES[name]()
        sid = O.SOLVER_BY_NAME[solver]
        sp = O.solver_params(*params) if solver == "acoustics" else O.solver_params(
            gamma=params[0], efix=bool(params[1]))
        out = os.path.join(HERE, f"claw2ez_{name}")
        shutil.rmtree(out, ignore_errors=True)
        os.makedirs(out)
        with tempfile.TemporaryDirectory() as tmp:
            rd.write(tmp)
            O.ref_claw2ez(tmp, sid, sp, q0)
            for f in sorted(os.listdir(tmp)):
                with open(os.path.join(tmp, f), "rb") as src, \
                        open(os.path.join(out, f + ".gz"), "wb") as raw, \
                        gzip.GzipFile(fileobj=raw, mode="wb", mtime=0, filename="") as dst:
                    dst.write(src.read())
        print(name, sorted(os.listdir(out)))


if __name__ == "__main__":
    main()
