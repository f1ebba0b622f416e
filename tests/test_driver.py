"""CPU checks of the run-data files, the frame files and the claw2ez driver
logic.

* ClawRunData -> claw.data -> read back in claw2ez's read order.
* Frame files: the Python writer against the reference out2 (compiled from
  src/2d/out2.f into oracle/_ref), byte for byte, including three-digit
  exponents and the |q| < 1e-99 flush; the Fortran Ew.d/Dw.d formatter
  against values whose printed form is fixed by the edit descriptor.
* Driver: classic_amd.driver with the oracle's claw2 as the stepper against
  the reference claw2ez run on the same claw.data, every output file byte
  for byte (output styles 1, 2 and 3, and a restart), and against the
  committed reference outputs where the reference build is absent.
"""
import gzip
import os

import numpy as np
import pytest

from classic_amd import driver, output, rundata
from oracle import oracle as O
from tests import ez_cases as E

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")


def _files(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))
            if f.startswith("fort.")}


def _golden_files(name):
    d = os.path.join(GOLDEN, f"claw2ez_{name}")
    out = {}
    for f in sorted(os.listdir(d)):
        if f.startswith("fort.") and f.endswith(".gz"):
            out[f[:-3]] = gzip.open(os.path.join(d, f)).read()
    return out


def _oracle_params(solver, params):
    if solver == "acoustics":
        return O.solver_params(*params)
    return O.solver_params(gamma=params[0], efix=bool(params[1]))


def _run_driver(tmp, name, restart_from=None):
    rd, q0, solver, params = E.CASES[name]()
    if restart_from is not None:
        rd.clawdata.restart = True
        rd.clawdata.restart_file = f"fort.q{restart_from:04d}"
    rd.write(tmp)
    cd = rd.clawdata
    sid = O.SOLVER_BY_NAME[solver]
    grid = O.OracleGrid2D(sid, cd.num_cells[0], cd.num_cells[1], _oracle_params(solver, params))
    return driver.claw2ez(tmp, qinit=E.qinit_from(q0), solver=solver, grid=grid), rd, q0


# ------------------------------------------------------------------ formats

@pytest.mark.parametrize("x,w,d,letter,text", [
    (1.0, 26, 16, "E", "    0.1000000000000000E+01"),
    (-2.5e-5, 26, 16, "E", "   -0.2500000000000000E-04"),
    (1e100, 26, 16, "E", "    0.1000000000000000+101"),
    (-1e-120, 26, 16, "E", "   -0.1000000000000000-119"),
    (0.0, 26, 16, "E", "    0.0000000000000000E+00"),
    (0.1, 26, 16, "E", "    0.1000000000000000E+00"),
    (0.09999999999999999, 26, 16, "E", "    0.9999999999999999E-01"),
    (9.9999999999999999e-1, 26, 16, "E", "    0.1000000000000000E+01"),
    (3.14159, 18, 8, "E", "    0.31415900E+01"),
    (0.0045, 15, 4, "D", "     0.4500D-02"),
    (0.99995, 15, 4, "D", "     0.1000D+01"),
])
def test_fortran_e(x, w, d, letter, text):
    assert output.fortran_e(x, w, d, letter) == text


def test_claw_data_roundtrip(tmp_path):
    rd, _, _, _ = E.euler_case(3)
    cd = rd.clawdata
    cd.bc_lower = ["periodic", "wall"]
    cd.bc_upper = ["periodic", "extrap"]
    cd.dimensional_split = "strang"
    rd.write(str(tmp_path))
    c = rundata.read_claw_data(str(tmp_path / "claw.data"))
    assert c.ndim == 2 and c.num_cells == [20, 16] and c.meqn == 4 and c.mwaves == 3
    assert c.lower == [0.0, 0.0] and c.upper == [1.0, 1.0]
    assert c.outstyle == 3 and c.nstepout == 2 and c.nout == 6
    assert c.method == [1, 2, -2, 0, 0, 0, 0]  # dimensional_split 2 -> method(3) = -2
    assert c.mthlim == [4, 4, 4]
    assert c.mthbc == [2, 2, 3, 1]  # (x lower, x upper, y lower, y upper)
    assert c.dtv[:2] == [0.02, 1e99] and c.cflv[:2] == [1.0, 0.9] and c.nv[0] == 5000
    assert c.mbc == 2 and c.restart is False
    text = open(tmp_path / "claw.data").read().splitlines()
    # opendatafile skips the comment lines and the one line after them
    assert text[0].startswith("#") and text[2] == ""


def test_claw_data_styles_and_names(tmp_path):
    rd, _, _, _ = E.acoustics_case(2)
    cd = rd.clawdata
    cd.limiter = ["superbee", "vanleer"]
    cd.restart = True
    cd.restart_file = "fort.chk00012"
    cd.num_aux = 2
    cd.output_aux_components = [1]
    cd.output_aux_onlyonce = False
    rd.write(str(tmp_path))
    c = rundata.read_claw_data(str(tmp_path / "claw.data"))
    assert c.outstyle == 2 and c.tout == [0.05, 0.12, 0.3] and c.nout == 3
    assert c.mthlim == [2, 3] and c.maux == 2 and c.method[6] == 2
    assert c.iout_aux == [1, 0] and c.outaux_always is True
    assert c.restart is True and c.restart_frame == 12
    u = rundata.read_user_data(str(tmp_path / "setprob.data"))
    assert [float(v[0]) for v in u] == [1.0, 4.0]
    with pytest.raises(ValueError):
        rd.clawdata.limiter = ["mc"]
        rd.clawdata.records()


def test_frame_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    q = rng.standard_normal((9, 11, 3))
    q[3, 4, 1] = 1e-120  # flushed to zero by the writer
    output.write_frame(str(tmp_path), 7, 0.25, q, 2, [0.0, -1.0], [0.1, 0.2])
    t, q2 = output.read_frame(str(tmp_path), 7, 2, 3)
    assert t == 0.25
    assert q[3, 4, 1] == 0.0
    # 16 significant digits (E26.16) keep a double to within one part in 1e16
    np.testing.assert_allclose(q2[2:-2, 2:-2], q[2:-2, 2:-2], rtol=6e-16, atol=0)
    assert output.read_frame_header(str(tmp_path), 7, 2) == ([7, 5], [0.0, -1.0], [0.1, 0.2])
    # writing what was read back reproduces the file
    first = open(tmp_path / "fort.q0007").read()
    output.write_frame(str(tmp_path), 7, 0.25, q2, 2, [0.0, -1.0], [0.1, 0.2])
    assert open(tmp_path / "fort.q0007").read() == first


@needs_ref
@pytest.mark.parametrize("outaux", [False, True])
def test_out2_matches_reference(tmp_path, outaux):
    rng = np.random.default_rng(1)
    q = rng.standard_normal((10, 13, 4)) * 10.0 ** rng.integers(-30, 30, (10, 13, 4))
    q[4, 5, 0] = 3e-120
    q[5, 6, 2] = -2.5e150
    aux = rng.uniform(0.5, 2.0, (10, 13, 2)) if outaux else None
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    qa = q.copy()
    O.ref_out2(str(a), qa, 0.375, 12, 2, [-1.0, 0.5], [2.0 / 9, 1.0 / 6], aux=aux, outaux=outaux)
    qb = q.copy()
    output.write_frame(str(b), 12, 0.375, qb, 2, [-1.0, 0.5], [2.0 / 9, 1.0 / 6], aux=aux,
                       write_aux=outaux)
    fa, fb = _files(str(a)), _files(str(b))
    assert sorted(fa) == sorted(fb)
    for f in fa:
        assert fa[f] == fb[f], f
    np.testing.assert_array_equal(qa, qb)  # the same in-place flush


@needs_ref
@pytest.mark.parametrize("name", list(E.CASES))
def test_driver_matches_reference_claw2ez(tmp_path, name):
    rd, q0, solver, params = E.CASES[name]()
    ref_dir = tmp_path / "ref"
    rd.write(str(ref_dir))
    O.ref_claw2ez(str(ref_dir), O.SOLVER_BY_NAME[solver], _oracle_params(solver, params), q0)
    mine = tmp_path / "mine"
    res, _, _ = _run_driver(str(mine), name)
    fr, fm = _files(str(ref_dir)), _files(str(mine))
    assert sorted(fr) == sorted(fm)
    for f in fr:
        assert fr[f] == fm[f], f
    assert res.info == 0 and len(res.frames) == len([f for f in fr if f.startswith("fort.q")])


@needs_ref
def test_driver_restart_matches_reference(tmp_path):
    """Restart from frame 1 (claw.data restart record): the later frames and
    fort.info equal the reference driver's restart run."""
    name = "acoustics_style1"
    rd, q0, solver, params = E.CASES[name]()
    base = tmp_path / "base"
    _run_driver(str(base), name)
    for sub in ("ref", "mine"):
        d = tmp_path / sub
        d.mkdir()
        for f in ("fort.q0001", "fort.t0001"):
            (d / f).write_bytes((base / f).read_bytes())
    rd.clawdata.restart = True
    rd.clawdata.restart_file = "fort.q0001"
    rd.write(str(tmp_path / "ref"))
    O.ref_claw2ez(str(tmp_path / "ref"), O.SOLVER_BY_NAME[solver], _oracle_params(solver, params),
                  q0)
    res, _, _ = _run_driver(str(tmp_path / "mine"), name, restart_from=1)
    fr, fm = _files(str(tmp_path / "ref")), _files(str(tmp_path / "mine"))
    assert sorted(fr) == sorted(fm)
    for f in fr:
        assert fr[f] == fm[f], f
    assert [f for f, _ in res.frames] == [2, 3]


@pytest.mark.parametrize("name", E.GOLDEN)
def test_driver_matches_committed_reference_outputs(tmp_path, name):
    """The same comparison against the reference outputs committed under
    tests/golden (written by tests/golden/make_golden_ez.py)."""
    _run_driver(str(tmp_path), name)
    fg, fm = _golden_files(name), _files(str(tmp_path))
    assert sorted(fg) == sorted(fm)
    for f in fg:
        assert fg[f] == fm[f], f


# --------------------------------------------------------------- 1D and 3D

from oracle import oracle1 as O1  # noqa: E402
from oracle import oracle3 as O3  # noqa: E402

needs_ref1 = pytest.mark.skipif(not O1.ref_available(), reason="reference build (oracle/_ref) absent")
needs_ref3 = pytest.mark.skipif(not O3.ref_available(), reason="reference build (oracle/_ref) absent")


def _run_driver1(tmp, name):
    rd, q0, solver, params = E.CASES1[name]()
    rd.write(tmp)
    grid = O1.OracleGrid1D(O1.SOLVER_ACOUSTICS, rd.clawdata.num_cells[0], O1.params1(*params))
    return driver.claw1ez(tmp, qinit=E.qinit_from(q0), solver=solver, grid=grid)


def _run_driver3(tmp, name):
    rd, (q0, aux0), solver, _ = E.CASES3[name]()
    rd.write(tmp)
    n = rd.clawdata.num_cells
    grid = O3.OracleGrid3D(n[0], n[1], n[2], lower=rd.clawdata.lower, upper=rd.clawdata.upper)
    return driver.claw3ez(tmp, qinit=E.qinit_from(q0), setaux=E.setaux_from(aux0), grid=grid)


def _same_dirs(a, b):
    fa, fb = _files(a), _files(b)
    assert sorted(fa) == sorted(fb)
    for f in fa:
        assert fa[f] == fb[f], f


@needs_ref1
@pytest.mark.parametrize("name", list(E.CASES1))
def test_driver1_matches_reference_claw1ez(tmp_path, name):
    rd, q0, solver, params = E.CASES1[name]()
    rd.write(str(tmp_path / "ref"))
    O1.ref_claw1ez(str(tmp_path / "ref"), O1.SOLVER_ACOUSTICS, O1.params1(*params), q0)
    res = _run_driver1(str(tmp_path / "mine"), name)
    _same_dirs(str(tmp_path / "ref"), str(tmp_path / "mine"))
    assert res.info == 0


@needs_ref3
@pytest.mark.parametrize("name", list(E.CASES3))
def test_driver3_matches_reference_claw3ez(tmp_path, name):
    rd, (q0, aux0), _, _ = E.CASES3[name]()
    rd.write(str(tmp_path / "ref"))
    O3.ref_claw3ez(str(tmp_path / "ref"), q0, aux0)
    res = _run_driver3(str(tmp_path / "mine"), name)
    _same_dirs(str(tmp_path / "ref"), str(tmp_path / "mine"))
    assert res.info == 0 and os.path.exists(tmp_path / "mine" / "fort.a0000")


@pytest.mark.parametrize("name", E.GOLDEN1 + E.GOLDEN3)
def test_driver13_matches_committed_reference_outputs(tmp_path, name):
    if name in E.CASES1:
        _run_driver1(str(tmp_path), name)
    else:
        _run_driver3(str(tmp_path), name)
    fg, fm = _golden_files(name), _files(str(tmp_path))
    assert sorted(fg) == sorted(fm)
    for f in fg:
        assert fg[f] == fm[f], f


def test_driver_user_routine_names(tmp_path):
    """The driver takes bc<d> / b4step<d> / src<d> for the run's dimension
    only, and a user boundary code (0) needs bc<d>."""
    name = E.GOLDEN1[0] if E.GOLDEN1 else list(E.CASES1)[0]
    rd, q0, solver, params = E.CASES1[name]()
    rd.write(str(tmp_path))
    grid = O1.OracleGrid1D(O1.SOLVER_ACOUSTICS, rd.clawdata.num_cells[0], O1.params1(*params))
    kw = dict(qinit=E.qinit_from(q0), solver=solver, grid=grid)
    with pytest.raises(ValueError, match="bc1, b4step1, src1"):
        driver.claw1ez(str(tmp_path), src2=lambda *a: None, **kw)
    with pytest.raises(TypeError, match="src4"):
        driver.claw1ez(str(tmp_path), src4=lambda *a: None, **kw)
    rd.clawdata.bc_lower = [0]
    rd.write(str(tmp_path))
    with pytest.raises(ValueError, match="needs the bc1 routine"):
        driver.claw1ez(str(tmp_path), **kw)
