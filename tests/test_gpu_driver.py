"""The claw2ez driver on the device against the reference driver's own
output files (tests/golden/claw2ez_<case>, written by the reference
src/2d/claw2ez.f through tests/golden/make_golden_ez.py): with EXACT
arithmetic every fort.qNNNN, fort.tNNNN and fort.info byte is the same."""
import gzip
import os

import numpy as np
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


def _golden(name):
    gdir = os.path.join(GOLDEN, f"claw2ez_{name}")
    return {f[:-3]: gzip.open(os.path.join(gdir, f)).read() for f in os.listdir(gdir)
            if f.startswith("fort.")}


def _mine(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in os.listdir(d)
            if f.startswith("fort.")}


@pytest.mark.parametrize("name", E.GOLDEN1)
def test_gpu_driver1_matches_reference_files(tmp_path, name):
    """1D: the device step1 is bit for bit with the reference, so every file
    is identical to the reference claw1ez's."""
    rd, q0, solver, params = E.CASES1[name]()
    rd.write(str(tmp_path))
    res = driver.claw1ez(str(tmp_path), qinit=E.qinit_from(q0), solver=solver,
                         solver_params=list(params))
    assert res.info == 0
    golden, mine = _golden(name), _mine(str(tmp_path))
    assert sorted(golden) == sorted(mine)
    for f in golden:
        assert golden[f] == mine[f], f


def _numbers(text):
    out = []
    for tok in text.decode().split():
        try:
            out.append(float(tok.replace("D", "E")))
        except ValueError:
            pass
    return np.array(out)


@pytest.mark.parametrize("name", E.GOLDEN3)
def test_gpu_driver3_matches_reference_files(tmp_path, name):
    """3D: the unsplit tile kernel sums each cell's increments in its own
    order (round-off differences), so the files are compared as numbers:
    every value within 1e-12 of the reference claw3ez's, relative to the
    largest value of its file."""
    rd, (q0, aux0), _, _ = E.CASES3[name]()
    rd.write(str(tmp_path))
    res = driver.claw3ez(str(tmp_path), qinit=E.qinit_from(q0), setaux=E.setaux_from(aux0))
    assert res.info == 0
    golden, mine = _golden(name), _mine(str(tmp_path))
    assert sorted(golden) == sorted(mine)
    for f in golden:
        a, b = _numbers(golden[f]), _numbers(mine[f])
        assert a.shape == b.shape, f
        scale = max(np.abs(a).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-12 * scale, f
