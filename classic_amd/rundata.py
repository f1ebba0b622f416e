"""Run parameters: the ``setrun.py`` interface of the reference's examples
and the ``claw.data`` / ``setprob.data`` files its drivers read.

``ClawRunData(claw_pkg, num_dim)`` with ``.clawdata`` (the attributes the
examples set, examples/acoustics_2d_radial/setrun.py:62-232),
``new_UserData(name, fname).add_param(name, value, descr)`` and ``write()``
produce the files; ``read_claw_data`` reads ``claw.data`` back in the order
of src/1d/claw1ez.f:44-150, src/2d/claw2ez.f:39-152 and
src/3d/claw3ez.f:51-162, which is what pins the file's content.

File layout: leading comment lines starting with '#', then one blank line
(the drivers' opendatafile, src/2d/opendatafile.f:31-40, consumes the first
line that is not a comment), then one record per read statement, the values
first and a '=: name' note after them, which a list-directed read ignores.
Logicals are written T/F; output_format is written as the number 1 (ascii),
since claw2ez reads it into an implicitly REAL variable (claw2ez.f:14, :76).
"""
from __future__ import annotations

import os
import re

LIMITERS = {"none": 0, "minmod": 1, "superbee": 2, "vanleer": 3, "van leer": 3, "mc": 4,
            "beamwarming": 5, "beam-warming": 5}
BCS = {"user": 0, "extrap": 1, "periodic": 2, "wall": 3}
SPLITS = {"unsplit": 0, "godunov": 1, "strang": 2}
TRANSVERSE = {"none": 0, "increment": 1, "all": 2}
SOURCES = {"none": 0, "godunov": 1, "strang": 2}
FORMATS = {"ascii": 1}


def _code(v, table, what):
    if isinstance(v, str):
        key = v.strip().lower()
        if key not in table:
            raise ValueError(f"unknown {what} {v!r}; expected one of {sorted(table)}")
        return table[key]
    return int(v)


class ClawInputData:
    """The ``clawdata`` attributes of the reference's setrun.py files."""

    def __init__(self, num_dim=2):
        self.num_dim = num_dim
        self.lower = [0.0] * num_dim
        self.upper = [1.0] * num_dim
        self.num_cells = [100] * num_dim
        self.num_eqn = 1
        self.num_waves = 1
        self.num_aux = 0
        self.capa_index = 0
        self.t0 = 0.0
        self.restart = False
        self.restart_file = ""
        self.output_style = 1
        self.num_output_times = 10
        self.tfinal = 1.0
        self.output_t0 = True
        self.output_times = []
        self.output_step_interval = 1
        self.total_steps = 1
        self.output_format = "ascii"
        self.output_q_components = "all"
        self.output_aux_components = "none"
        self.output_aux_onlyonce = True
        self.verbosity = 0
        self.dt_variable = True
        self.dt_initial = 1e-5
        self.dt_max = 1e99
        self.cfl_desired = 0.9
        self.cfl_max = 1.0
        self.steps_max = 50000
        self.order = 2
        self.dimensional_split = "unsplit"
        self.transverse_waves = 2
        self.limiter = [4]
        self.use_fwaves = False
        self.source_split = 0
        self.num_ghost = 2
        self.bc_lower = ["extrap"] * num_dim
        self.bc_upper = ["extrap"] * num_dim

    # ---------------------------------------------------------------- values
    def records(self):
        """(values, name) records in the drivers' read order."""
        d = self.num_dim
        L = lambda b: "T" if b else "F"  # noqa: E731
        recs = [([d], "num_dim"), (list(self.lower), "lower"), (list(self.upper), "upper"),
                ([int(n) for n in self.num_cells], "num_cells"), ([self.num_eqn], "num_eqn"),
                ([self.num_waves], "num_waves"), ([self.num_aux], "num_aux"),
                ([float(self.t0)], "t0"), ([int(self.output_style)], "output_style")]
        if self.output_style == 1:
            recs += [([int(self.num_output_times)], "num_output_times"),
                     ([float(self.tfinal)], "tfinal"), ([L(self.output_t0)], "output_t0")]
        elif self.output_style == 2:
            times = [float(t) for t in self.output_times]
            recs += [([len(times)], "num_output_times"), (times, "output_times")]
        elif self.output_style == 3:
            recs += [([int(self.output_step_interval)], "output_step_interval"),
                     ([int(self.total_steps)], "total_steps"), ([L(self.output_t0)], "output_t0")]
        else:
            raise ValueError("output_style must be 1, 2 or 3")
        recs.append(([_code(self.output_format, FORMATS, "output_format")], "output_format"))
        recs.append((self._components(self.output_q_components, self.num_eqn), "output_q_components"))
        if self.num_aux > 0:
            recs.append((self._components(self.output_aux_components, self.num_aux),
                         "output_aux_components"))
            recs.append(([L(self.output_aux_onlyonce)], "output_aux_onlyonce"))
        recs += [([float(self.dt_initial)], "dt_initial"), ([float(self.dt_max)], "dt_max"),
                 ([float(self.cfl_max)], "cfl_max"), ([float(self.cfl_desired)], "cfl_desired"),
                 ([int(self.steps_max)], "steps_max"), ([L(self.dt_variable)], "dt_variable"),
                 ([int(self.order)], "order")]
        if d > 1:
            recs.append(([self._transverse()], "transverse_waves"))
            recs.append(([_code(self.dimensional_split, SPLITS, "dimensional_split")],
                         "dimensional_split"))
        recs += [([int(self.verbosity)], "verbosity"),
                 ([_code(self.source_split, SOURCES, "source_split")], "source_split"),
                 ([int(self.capa_index)], "capa_index"), ([L(self.use_fwaves)], "use_fwaves")]
        lim = list(self.limiter) if isinstance(self.limiter, (list, tuple)) else [self.limiter]
        if len(lim) != self.num_waves:
            raise ValueError("limiter needs one entry per wave (len(limiter) == num_waves)")
        recs.append(([_code(v, LIMITERS, "limiter") for v in lim], "limiter"))
        recs.append(([int(self.num_ghost)], "num_ghost"))
        recs.append(([_code(v, BCS, "bc") for v in self.bc_lower], "bc_lower"))
        recs.append(([_code(v, BCS, "bc") for v in self.bc_upper], "bc_upper"))
        recs.append(([L(self.restart)], "restart"))
        if self.restart:
            recs.append(([restart_frame(self.restart_file)], "restart_frame"))
        return recs

    def _transverse(self):
        return _code(self.transverse_waves, TRANSVERSE, "transverse_waves")

    @staticmethod
    def _components(spec, n):
        if isinstance(spec, str):
            s = spec.strip().lower()
            if s == "all":
                return [1] * n
            if s == "none":
                return [0] * n
            raise ValueError(f"unknown component selection {spec!r}")
        flags = [0] * n
        for c in spec:  # 1-based component numbers, as in the reference's setrun files
            flags[int(c) - 1] = 1
        return flags

    def write(self, path):
        with open(path, "w") as f:
            f.write("# Clawpack Classic run parameters (claw.data), read by the claw*ez drivers\n")
            f.write("# in the order of their read statements.\n")
            f.write("\n")
            for vals, name in self.records():
                text = " ".join(_fmt(v) for v in vals)
                f.write(f"{text:<30s} =: {name}\n")


def _fmt(v):
    if isinstance(v, str):
        return v
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, (int,)) and not isinstance(v, bool):
        return str(v)
    return repr(float(v))


def restart_frame(restart_file):
    """Frame number of a restart file name (the digits at its end, e.g.
    fort.q0006 or fort.chk00006 -> 6)."""
    m = re.search(r"(\d+)\s*$", str(restart_file))
    if not m:
        raise ValueError(f"cannot read a frame number from restart_file {restart_file!r}")
    return int(m.group(1))


class UserData:
    """Problem parameters written to their own data file (setprob.data)."""

    def __init__(self, name, fname):
        self.name = name
        self.fname = fname
        self.params = []

    def add_param(self, name, value, descr=""):
        self.params.append((name, value, descr))
        setattr(self, name, value)

    def __getitem__(self, name):
        for n, v, _ in self.params:
            if n == name:
                return v
        raise KeyError(name)

    def write(self, path):
        with open(path, "w") as f:
            f.write(f"# {self.name} parameters\n")
            f.write("\n")
            for name, value, descr in self.params:
                if isinstance(value, (list, tuple)):
                    text = " ".join(_fmt(v) for v in value)
                else:
                    text = _fmt(value)
                f.write(f"{text:<30s} =: {name}   {descr}\n")


class ClawRunData:
    """Run data of one Clawpack Classic run (``claw_pkg`` must be "classic")."""

    def __init__(self, claw_pkg="classic", num_dim=2):
        if str(claw_pkg).lower() != "classic":
            raise ValueError("Expected claw_pkg = 'classic'")
        if num_dim not in (1, 2, 3):
            raise ValueError("num_dim must be 1, 2 or 3")
        self.claw_pkg = "classic"
        self.num_dim = num_dim
        self.clawdata = ClawInputData(num_dim)
        self.user_data = []

    def new_UserData(self, name, fname):  # noqa: N802 (the reference's name)
        u = UserData(name, fname)
        self.user_data.append(u)
        setattr(self, name, u)
        return u

    def write(self, out_dir="."):
        os.makedirs(out_dir, exist_ok=True)
        self.clawdata.write(os.path.join(out_dir, "claw.data"))
        for u in self.user_data:
            u.write(os.path.join(out_dir, u.fname))


# --------------------------------------------------------------------- reading

def _records(path):
    """Data records of a claw*ez data file: comment lines are skipped and so
    is the first line after them (opendatafile.f:31-40)."""
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines) and lines[k][:1] == "#":
        k += 1
    return lines[k + 1:]


def _tokens(line):
    body = line.split("=:")[0]
    return [t for t in re.split(r"[,\s]+", body.strip()) if t]


def _float(t):
    return float(t.lower().replace("d", "e"))


def _int(t):
    return int(float(t.lower().replace("d", "e"))) if re.search(r"[.eEdD]", t) else int(t)


def _logical(t):
    s = t.strip().lstrip(".").lower()
    if s.startswith("t"):
        return True
    if s.startswith("f"):
        return False
    raise ValueError(f"not a Fortran logical: {t!r}")


class ClawConfig(dict):
    """claw.data as the drivers read it: mx/my/mz, lower/upper, method(1:7),
    dtv, cflv, nv, mthlim, mthbc (reference order: lower/upper per
    dimension, mthbc(1)=x lower, (2)=x upper, ...), output settings, restart
    flag and frame."""

    __getattr__ = dict.__getitem__


def read_claw_data(path):
    """Read claw.data in the order of claw1ez/claw2ez/claw3ez."""
    recs = iter(_records(path))
    nxt = lambda: _tokens(next(recs))  # noqa: E731
    c = ClawConfig()
    c["ndim"] = nd = _int(nxt()[0])
    c["lower"] = [_float(t) for t in nxt()[:nd]]
    c["upper"] = [_float(t) for t in nxt()[:nd]]
    c["num_cells"] = [_int(t) for t in nxt()[:nd]]
    c["meqn"] = _int(nxt()[0])
    c["mwaves"] = _int(nxt()[0])
    c["maux"] = maux = _int(nxt()[0])
    c["t0"] = _float(nxt()[0])
    c["outstyle"] = os_ = _int(nxt()[0])
    c["tout"] = []
    c["nstepout"] = 1
    if os_ == 1:
        c["nout"] = _int(nxt()[0])
        c["tfinal"] = _float(nxt()[0])
        c["output_t0"] = _logical(nxt()[0])
    elif os_ == 2:
        c["nout"] = n = _int(nxt()[0])
        toks = []
        while len(toks) < n:
            toks += nxt()
        c["tout"] = [_float(t) for t in toks[:n]]
    elif os_ == 3:
        c["nstepout"] = _int(nxt()[0])
        c["nout"] = _int(nxt()[0])
        c["output_t0"] = _logical(nxt()[0])
    else:
        raise ValueError(f"unrecognized output style {os_}")
    c["output_format"] = _float(nxt()[0])
    c["iout_q"] = [_int(t) for t in nxt()[: c["meqn"]]]
    c["outaux_init_only"] = c["outaux_always"] = False
    if maux > 0:
        c["iout_aux"] = iout_aux = [_int(t) for t in nxt()[:maux]]
        init_only = _logical(nxt()[0])
        if any(v != 0 for v in iout_aux):
            c["outaux_init_only"] = init_only
            c["outaux_always"] = not init_only
    dtv1 = _float(nxt()[0])
    dtv2 = _float(nxt()[0])
    cflv1 = _float(nxt()[0])
    cflv2 = _float(nxt()[0])
    c["dtv"] = [dtv1, dtv2, 0.0, 0.0, 0.0]
    c["cflv"] = [cflv1, cflv2, 0.0, 0.0]
    c["nv"] = [_int(nxt()[0]), 0]
    method = [0] * 7
    method[0] = 1 if _logical(nxt()[0]) else 0
    method[1] = _int(nxt()[0])
    if nd > 1:
        method[2] = _int(nxt()[0])
        split = _int(nxt()[0])
        if split > 0:
            method[2] = -split
    method[3] = _int(nxt()[0])
    method[4] = _int(nxt()[0])
    method[5] = _int(nxt()[0])
    method[6] = maux
    c["method"] = method
    c["use_fwaves"] = _logical(nxt()[0])
    c["mthlim"] = [_int(t) for t in nxt()[: c["mwaves"]]]
    c["mbc"] = _int(nxt()[0])
    lo = [_int(t) for t in nxt()[:nd]]
    hi = [_int(t) for t in nxt()[:nd]]
    mthbc = []
    for k in range(nd):
        mthbc += [lo[k], hi[k]]
    c["mthbc"] = mthbc
    c["restart"] = False
    c["restart_frame"] = 0
    try:
        c["restart"] = _logical(nxt()[0])
        if c["restart"]:
            c["restart_frame"] = _int(nxt()[0])
    except (StopIteration, IndexError, ValueError):
        # claw2ez.f:142: a missing or unreadable restart record means no restart
        c["restart"] = False
    return c


def read_user_data(path):
    """Values of a setprob-style data file, in order (strings for each
    record's first token list)."""
    out = []
    for line in _records(path):
        toks = _tokens(line)
        if toks:
            out.append(toks)
    return out
