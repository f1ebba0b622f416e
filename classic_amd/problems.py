"""Initial data for the benchmark and test problems (the ``qinit`` role of
the reference's plug-in contract, src/2d/claw2ez.f:244-245).

Arrays are returned in the reference layout q(meqn, 1-mbc:mx+mbc,
1-mbc:my+mbc), i.e. C-contiguous numpy arrays of shape (my+2mbc, mx+2mbc,
meqn); ghost cells are left at zero (bc2 fills them before every step).
"""
from __future__ import annotations

import numpy as np


def cell_centers(lower, upper, n, mbc=2):
    d = (upper - lower) / n
    idx = np.arange(1 - mbc, n + mbc + 1)
    return lower + (idx - 0.5) * d, d


def acoustics_pulse(mx=200, my=200, mbc=2, lower=(-1.0, -1.0), upper=(1.0, 1.0),
                    width=0.35):
    """2D acoustics, q = (p, u, v): p = 1 + cos(pi r / width) for r < width,
    else p = 1, with r the distance of the cell centre from the origin;
    u = v = 0.  (The literal formula jumps from 0 to 1 at r = width.)"""
    xc, dx = cell_centers(lower[0], upper[0], mx, mbc)
    yc, dy = cell_centers(lower[1], upper[1], my, mbc)
    q = np.zeros((my + 2 * mbc, mx + 2 * mbc, 3))
    X, Y = np.meshgrid(xc, yc)
    r = np.sqrt(X * X + Y * Y)
    p = np.where(r < width, 1.0 + np.cos(np.pi * r / width), 1.0)
    q[..., 0] = p
    q[:mbc, :, :] = 0.0
    q[-mbc:, :, :] = 0.0
    q[:, :mbc, :] = 0.0
    q[:, -mbc:, :] = 0.0
    return q, dx, dy


def euler_quadrant_states(seed):
    """Per-quadrant primitive states (rho, u, v, p) drawn from
    numpy.random.default_rng(seed) in the fixed order NE, NW, SW, SE, each
    as rho ~ U[0.1, 2], u ~ U[-1, 1], v ~ U[-1, 1], p ~ U[0.1, 2]."""
    rng = np.random.default_rng(seed)
    states = []
    for _ in range(4):
        rho = rng.uniform(0.1, 2.0)
        u = rng.uniform(-1.0, 1.0)
        v = rng.uniform(-1.0, 1.0)
        p = rng.uniform(0.1, 2.0)
        states.append((rho, u, v, p))
    return states


def euler_quadrants(mx=4096, my=4096, seed=0, gamma=1.4, mbc=2, split=(0.5, 0.5),
                    lower=(0.0, 0.0), upper=(1.0, 1.0), states=None):
    """2D Euler four-quadrant Riemann problem on [0,1]^2 split at (0.5, 0.5);
    q = (rho, rho u, rho v, E) with E = p/(gamma-1) + rho (u^2+v^2)/2."""
    xc, dx = cell_centers(lower[0], upper[0], mx, mbc)
    yc, dy = cell_centers(lower[1], upper[1], my, mbc)
    if states is None:
        states = euler_quadrant_states(seed)
    q = np.zeros((my + 2 * mbc, mx + 2 * mbc, 4))
    east = xc >= split[0]
    north = yc >= split[1]
    for k, (rho, u, v, p) in enumerate(states):
        ex = east if k in (0, 3) else ~east
        ny = north if k in (0, 1) else ~north
        mask = ny[:, None] & ex[None, :]
        q[mask, 0] = rho
        q[mask, 1] = rho * u
        q[mask, 2] = rho * v
        q[mask, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return q, dx, dy


def layered_medium_aux(mx, my, mz, lower=(-1.0, -1.0, -1.0), upper=(1.0, 1.0, 1.0), mbc=2,
                       below=(2.0, 2.0), above=(1.0, 1.0)):
    """aux = (Z, c) of a two-layer medium split at z = 0, ghost cells
    included: `below` for cell centres with z < 0, `above` otherwise (the
    setaux of examples/acoustics_3d_heterogeneous/setaux.f:19-35).  Shape
    (mz+2mbc, my+2mbc, mx+2mbc, 2)."""
    zc, _ = cell_centers(lower[2], upper[2], mz, mbc)
    aux = np.empty((mz + 2 * mbc, my + 2 * mbc, mx + 2 * mbc, 2))
    lo = zc < 0.0
    aux[lo] = below
    aux[~lo] = above
    return aux


def acoustics3d_heterogeneous(n=20, mbc=2):
    """The heterogeneous-acoustics example (examples/acoustics_3d_heterogeneous):
    [-1,1]^3, q = 0, Z = c = 2 below z = 0 and Z = c = 1 above, driven by the
    piston boundary condition at the top (mthbc(6) = 0).  Returns (q, aux)."""
    q = np.zeros((n + 2 * mbc,) * 3 + (4,))
    return q, layered_medium_aux(n, n, n, mbc=mbc)


def acoustics3d_pulse(n=512, seed=0, mbc=2, width=0.01, center=None):
    """3D acoustics on [-1,1]^3, q = (p, u, v, w): p = exp(-|x - x0|^2 / width^2)
    with x0 drawn from default_rng(seed).uniform(-0.5, 0.5, 3) (or the given
    center), u = v = w = 0; aux = (Z, c): rho = 1, c = 1 (Z = 1) for z < 0
    and rho = 3, c = 2 (Z = 6) above.  Returns (q, aux, x0)."""
    if center is None:
        x0 = np.random.default_rng(seed).uniform(-0.5, 0.5, 3)
    else:
        x0 = np.asarray(center, dtype=np.float64)
    xc, _ = cell_centers(-1.0, 1.0, n, mbc)
    q = np.zeros((n + 2 * mbc,) * 3 + (4,))
    ex = np.exp(-((xc - x0[0]) ** 2) / width ** 2)
    ey = np.exp(-((xc - x0[1]) ** 2) / width ** 2)
    ez = np.exp(-((xc - x0[2]) ** 2) / width ** 2)
    q[..., 0] = ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    aux = layered_medium_aux(n, n, n, mbc=mbc, below=(1.0, 1.0), above=(6.0, 2.0))
    return q, aux, x0


def acoustics1d_gaussian(mx=400, mbc=2, lower=-1.0, upper=1.0, beta=200.0):
    """examples/acoustics_1d_example1: q = (p, u) with p = exp(-beta (x-0.3)^2),
    u = 0 (qinit.f90).  Returns (q, dx), q of shape (mx+2mbc, 2)."""
    xc, dx = cell_centers(lower, upper, mx, mbc)
    q = np.zeros((mx + 2 * mbc, 2))
    q[mbc:-mbc, 0] = np.exp(-beta * (xc[mbc:-mbc] - 0.3) ** 2)
    return q, dx


def acoustics1d_heterogeneous(mx=500, mbc=2, lower=-5.0, upper=5.0, rhol=1.0, cl=1.0, rhor=4.0,
                              cr=0.5):
    """examples/acoustics_1d_heterogeneous with ic = 1: p = u = sqrt(1-(x+3)^2)
    on -4 < x < -2 (qinit.f); aux = (Z, c) = (rho*c, c) left / right of x = 0
    (setaux.f, setprob.f).  Returns (q, aux, dx)."""
    xc, dx = cell_centers(lower, upper, mx, mbc)
    q = np.zeros((mx + 2 * mbc, 2))
    xi = xc[mbc:-mbc]
    hump = np.where((xi > -4.0) & (xi < -2.0), np.sqrt(np.maximum(1.0 - (xi + 3.0) ** 2, 0.0)), 0.0)
    q[mbc:-mbc, 0] = hump
    q[mbc:-mbc, 1] = hump
    aux = np.empty((mx + 2 * mbc, 2))
    left = xc < 0.0
    aux[left] = (rhol * cl, cl)
    aux[~left] = (rhor * cr, cr)
    return q, aux, dx


def annulus_hump(mx=40, my=120, mbc=2, xlower=0.2, xupper=1.0, ylower=0.0,
                 yupper=2.0 * np.pi, A=(1.0, -1.0), beta=(40.0, 40.0), centre=((-0.5, 0.0),
                                                                              (0.5, 0.0))):
    """examples/advection_2d_annulus qinit.f: two Gaussian humps
    A_k exp(-beta_k |p - c_k|^2) at the mapped cell centres p = (r cos th,
    r sin th) of the (r, th) grid.  Returns q of shape (my+2mbc, mx+2mbc, 1)."""
    xc, _ = cell_centers(xlower, xupper, mx, mbc)
    yc, _ = cell_centers(ylower, yupper, my, mbc)
    R, T = np.meshgrid(xc, yc)
    xp, yp = R * np.cos(T), R * np.sin(T)
    q = np.zeros((my + 2 * mbc, mx + 2 * mbc, 1))
    v = sum(a * np.exp(-b * ((xp - c[0]) ** 2 + (yp - c[1]) ** 2))
            for a, b, c in zip(A, beta, centre))
    q[mbc:-mbc, mbc:-mbc, 0] = v[mbc:-mbc, mbc:-mbc]
    return q


def batch_problem_specs(nprob=256, seed=0, side_lo=128, side_hi=1024, p_euler=0.5):
    """Specs of a batch of independent 2D problems: each grid side drawn
    log-uniformly in [side_lo, side_hi] (mx and my independently), each
    problem Euler (four random quadrant states, extrapolation boundaries)
    with probability p_euler, else acoustics with rho ~ U[0.5, 4] and
    bulk ~ U[1, 8] (the radial pressure pulse of ``acoustics_pulse``, walls).
    Returns a list of dicts with solver ("euler"/"acoustics"), mx, my,
    params, mthbc and a per-problem seed."""
    rng = np.random.default_rng(seed)
    specs = []
    lo, hi = np.log(side_lo), np.log(side_hi)
    for k in range(nprob):
        mx = int(round(np.exp(rng.uniform(lo, hi))))
        my = int(round(np.exp(rng.uniform(lo, hi))))
        if rng.uniform() < p_euler:
            specs.append(dict(solver="euler", mx=mx, my=my, params=(1.4, 1.0), mthbc=(1, 1, 1, 1),
                              seed=int(rng.integers(0, 2**31))))
        else:
            rho = rng.uniform(0.5, 4.0)
            bulk = rng.uniform(1.0, 8.0)
            specs.append(dict(solver="acoustics", mx=mx, my=my, params=(rho, bulk),
                              mthbc=(3, 3, 3, 3), seed=int(rng.integers(0, 2**31))))
    return specs


def batch_problem_q(spec, mbc=2):
    """Initial data (q, dx, dy) of one batch problem spec."""
    if spec["solver"] == "euler":
        return euler_quadrants(spec["mx"], spec["my"], seed=spec["seed"], mbc=mbc)
    return acoustics_pulse(spec["mx"], spec["my"], mbc=mbc)
