"""The claw2ez driver on the device against the reference driver's own
output files (tests/golden/claw2ez_<case>, written by the reference
src/2d/claw2ez.f through tests/golden/make_golden_ez.py): with EXACT
arithmetic every fort.qNNNN, fort.tNNNN and fort.info byte is the same."""
import gzip
import os

import pytest

from classic_amd import driver
from tests import ez_cases as E

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", E.GOLDEN)
def test_gpu_driver_matches_reference_files(tmp_path, name):
    rd, q0, solver, params = E.CASES[name]()
    rd.write(str(tmp_path))
    res = driver.claw2ez(str(tmp_path), qinit=E.qinit_from(q0), solver=solver,
                         solver_params=list(params), arith=0)
    assert res.info == 0
    gdir = os.path.join(GOLDEN, f"claw2ez_{name}")
    golden = {f[:-3]: gzip.open(os.path.join(gdir, f)).read() for f in os.listdir(gdir)
              if f.startswith("fort.")}
    mine = {f: open(os.path.join(tmp_path, f), "rb").read() for f in os.listdir(tmp_path)
            if f.startswith("fort.")}
    assert sorted(golden) == sorted(mine)
    for f in golden:
        assert golden[f] == mine[f], f
