"""Frame files of the reference's output routines: fort.qNNNN (solution),
fort.tNNNN (time and sizes), fort.aNNNN (aux arrays) and the fort.info
record of the claw*ez drivers.

Written exactly as src/1d/out1.f:25-110, src/2d/out2.f:24-128 and
src/3d/out3.f:40-131 write them (Fortran edit descriptors restated in
``fortran_e``), including the side effect of those routines: every value
below 1e-99 in magnitude is set to zero in the array being written, so a
driver that continues from that array continues from the flushed values.
The readers accept what these writers (and the reference) produce, for
restarts as in src/2d/restart2.f:3-66.
"""
from __future__ import annotations

import os

import numpy as np


def fortran_e(x: float, w: int, d: int, letter: str = "E") -> str:
    """x in the Fortran edit descriptor Ew.d (or Dw.d with letter="D"):
    0.<d digits><letter><sign><2 digits>, with the exponent letter dropped
    for a three-digit exponent, right-justified in w characters (the
    optional leading zero is dropped when it does not fit, and the field is
    w asterisks when nothing fits)."""
    x = float(x)
    if x != x:
        s = "NaN"
    elif x in (float("inf"), float("-inf")):
        s = ("-" if x < 0 else "") + "Infinity"
    else:
        neg = x < 0 or (x == 0.0 and str(x).startswith("-"))
        a = abs(x)
        if a == 0.0:
            digits, e = "0" * d, 0
        else:
            m = f"{a:.{d - 1}e}"  # d significant digits, correctly rounded
            mant, ex = m.split("e")
            digits = mant.replace(".", "")
            e = int(ex) + 1
        if abs(e) <= 99:
            exp = f"{letter}{'-' if e < 0 else '+'}{abs(e):02d}"
        elif abs(e) <= 999:
            exp = f"{'-' if e < 0 else '+'}{abs(e):03d}"
        else:
            return "*" * w
        body = "0." + digits + exp
        s = ("-" if neg else "") + body
        if len(s) > w:
            s = ("-" if neg else "") + body[1:]
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def _i5(n: int) -> str:
    s = str(int(n))
    return s.rjust(5) if len(s) <= 5 else "*****"


def frame_names(frame: int):
    """fort.q/t/a names of a frame number (four digits, out2.f:33-41)."""
    tag = f"{frame % 10000:04d}"
    return f"fort.q{tag}", f"fort.t{tag}", f"fort.a{tag}"


def flush_tiny(a):
    """The in-place zeroing of the output routines: |value| < 1e-99 -> 0."""
    a[np.abs(a) < 1e-99] = 0.0
    return a


_LABELS = {1: ["x"], 2: ["x", "y"], 3: ["x", "y", "z"]}
_SIZES = {1: ["mx"], 2: ["mx", "my"], 3: ["mx", "my", "mz"]}


def _header(ndim, ncells, lower, delta):
    lines = [_i5(1) + "                 grid_number", _i5(1) + "                 AMR_level"]
    for name, n in zip(_SIZES[ndim], ncells):
        lines.append(_i5(n) + "                 " + name)
    if ndim == 1:
        # out1.f:62-64 writes the grid in e18.8
        lines.append(fortran_e(lower[0], 18, 8) + "    xlow")
        lines.append(fortran_e(delta[0], 18, 8) + "    dx")
    else:
        for c, v in zip(_LABELS[ndim], lower):
            lines.append(fortran_e(v, 26, 16) + f"    {c}low")
        for c, v in zip(_LABELS[ndim], delta):
            lines.append(fortran_e(v, 26, 16) + f"    d{c}")
    lines.append("")
    return lines


def _row(vals):
    return "".join(fortran_e(v, 26, 16) for v in vals)


def _interior_cells(arr, ndim, mbc):
    """Interior cell values in the reference's loop order (x fastest; the
    arrays here are (..., y, x, m) with z slowest)."""
    s = slice(mbc, -mbc) if mbc > 0 else slice(None)
    return arr[(s,) * ndim]


def _body(interior, ndim):
    """Cell lines with the blank records of out1/out2/out3: one blank line
    (list-directed ' ', written as two spaces) after every x row, another
    after every y plane in 3D, and one at the end."""
    lines = []
    if ndim == 1:
        for cell in interior:
            lines.append(_row(cell))
        lines += ["  ", "  "]
    elif ndim == 2:
        for row in interior:
            for cell in row:
                lines.append(_row(cell))
            lines.append("  ")
        lines.append("  ")
    else:
        for plane in interior:
            for row in plane:
                for cell in row:
                    lines.append(_row(cell))
                lines.append("  ")
            lines.append("  ")
        lines.append("  ")
    return lines


def write_frame(outdir, frame, t, q, mbc, lower, delta, aux=None, write_aux=False):
    """Write fort.qNNNN and fort.tNNNN (and fort.aNNNN when write_aux) for
    the single grid q of shape (ny.., nx, meqn) in the reference layout with
    mbc ghost cells; flushes |q| < 1e-99 (and aux) in place as the
    reference does.  Returns the file names written."""
    ndim = q.ndim - 1
    meqn = q.shape[-1]
    ncells = [q.shape[ndim - 1 - k] - 2 * mbc for k in range(ndim)]
    maux = 0 if aux is None else aux.shape[-1]
    fq, ft, fa = frame_names(frame)
    os.makedirs(outdir, exist_ok=True)
    inner = _interior_cells(q, ndim, mbc)
    flush_tiny(inner)
    lines = _header(ndim, ncells, lower, delta) + _body(inner, ndim)
    with open(os.path.join(outdir, fq), "w") as f:
        f.write("\n".join(lines) + "\n")
    written = [fq, ft]
    if write_aux and aux is not None and maux > 0:
        ainner = _interior_cells(aux, ndim, mbc)
        flush_tiny(ainner)
        alines = _header(ndim, ncells, lower, delta) + _body(ainner, ndim)
        if ndim == 1:
            alines = alines[:-1]  # out1.f writes one closing blank record for aux
        with open(os.path.join(outdir, fa), "w") as f:
            f.write("\n".join(alines) + "\n")
        written.append(fa)
    tl = [fortran_e(t, 26, 16) + "    time",
          _i5(meqn) + "                 meqn",
          _i5(1) + "                 ngrids",
          _i5(maux) + "                 maux",
          _i5(ndim) + "                 ndim",
          _i5(mbc) + "                 nghost",
          "ascii".ljust(8).rjust(10) + "                 format",
          "", ""]
    with open(os.path.join(outdir, ft), "w") as f:
        f.write("\n".join(tl) + "\n")
    return written


def read_frame(outdir, frame, ndim, meqn, mbc=2):
    """Read fort.qNNNN / fort.tNNNN: returns (t, q) with q in the reference
    layout (ghost cells zero), as restart2.f reads a restart frame."""
    fq, ft, _ = frame_names(frame)
    with open(os.path.join(outdir, fq)) as f:
        toks = f.read().split()
    # header: grid_number, AMR_level, sizes, lows, deltas (value + label each)
    vals = toks[0::2][: 2 + 3 * ndim]
    ncells = [int(v) for v in vals[2: 2 + ndim]]
    nhead = 2 * (2 + 3 * ndim)
    data = np.array([float(v.replace("D", "E").replace("d", "e")) for v in toks[nhead:]])
    shape = tuple(reversed(ncells)) + (meqn,)
    q = np.zeros(tuple(n + 2 * mbc for n in reversed(ncells)) + (meqn,))
    q[(slice(mbc, -mbc),) * ndim] = data.reshape(shape)
    with open(os.path.join(outdir, ft)) as f:
        t = float(f.read().split()[0].replace("D", "E"))
    return t, q


def read_frame_header(outdir, frame, ndim):
    """(ncells, lower, delta) from a fort.qNNNN header."""
    fq, _, _ = frame_names(frame)
    with open(os.path.join(outdir, fq)) as f:
        toks = f.read().split()
    vals = toks[0::2][: 2 + 3 * ndim]
    ncells = [int(v) for v in vals[2: 2 + ndim]]
    lower = [float(v) for v in vals[2 + ndim: 2 + 2 * ndim]]
    delta = [float(v) for v in vals[2 + 2 * ndim: 2 + 3 * ndim]]
    return ncells, lower, delta


def info_record(tend, info, dtv, cflv, nv):
    """The fort.info record of claw2ez.f:304-307 (format 1010)."""
    steps = int(nv[1])
    st = str(steps)
    st = st.rjust(4) if len(st) <= 4 else "****"
    return ("tend =" + fortran_e(tend, 15, 4, "D") + "\n"
            + "info =" + str(int(info)).rjust(5) + "\n"
            + "smallest dt =" + fortran_e(dtv[2], 15, 4, "D") + "\n"
            + "largest dt =" + fortran_e(dtv[3], 15, 4, "D") + "\n"
            + "last dt =" + fortran_e(dtv[4], 15, 4, "D") + "\n"
            + "largest cfl =" + fortran_e(cflv[2], 15, 4, "D") + "\n"
            + "last cfl =" + fortran_e(cflv[3], 15, 4, "D") + "\n"
            + "steps taken =" + st + "\n"
            + "\n")
