"""Shared helpers for the parity tests (comparison bars and problem setups)."""
import numpy as np

from classic_amd import problems as P


def interior(q, mbc=2):
    return q[mbc:-mbc, mbc:-mbc]


FIELD_FLOOR = 1e-3


def max_rel_err(a, b, mbc=2):
    """Max relative error per field over the interior, maximised over fields:
    max|a_m - b_m| / scale_m with scale_m = max|b_m|.  A field that stays
    below FIELD_FLOOR times the largest field (e.g. the transverse momentum
    of a one-dimensional problem, which is zero up to round-off) has no
    scale of its own; it is measured against FIELD_FLOOR times the largest
    field instead, otherwise the ratio of two round-off noises is reported."""
    a = interior(a, mbc)
    b = interior(b, mbc)
    top = max(np.abs(b[..., m]).max() for m in range(b.shape[-1]))
    worst = 0.0
    for m in range(a.shape[-1]):
        scale = max(np.abs(b[..., m]).max(), FIELD_FLOOR * top)
        d = np.abs(a[..., m] - b[..., m]).max()
        worst = max(worst, d / scale if scale > 0 else d)
    return worst


CELL_FLOOR = 1e-3


def cell_rel_errs(a, b, mbc=2, floor=CELL_FLOOR):
    """Relative error cell by cell over the interior, per field:
    max |a - b| / max(|b|, floor * scale_m), scale_m as in max_rel_err.
    Unlike max_rel_err, a cell far below its field's largest value (a
    low-density corner, the tail of a pulse) is held to its own magnitude
    down to floor times the field's scale; floor = 0 holds every cell to
    itself (meaningful for fields bounded away from zero: density, energy,
    the pressure of a pulse on a background)."""
    nd = 3 if a.ndim == 4 else 2
    sl = (slice(mbc, -mbc),) * nd
    a = a[sl]
    b = b[sl]
    top = max(np.abs(b[..., m]).max() for m in range(b.shape[-1]))
    out = []
    for m in range(a.shape[-1]):
        scale = max(np.abs(b[..., m]).max(), FIELD_FLOOR * top)
        den = np.maximum(np.abs(b[..., m]), floor * scale)
        out.append(float((np.abs(a[..., m] - b[..., m]) / np.where(den > 0, den, 1.0)).max()))
    return out


def smooth_acoustics(mx, my, seed=0, mbc=2):
    rng = np.random.default_rng(seed)
    q = np.zeros((my + 2 * mbc, mx + 2 * mbc, 3))
    y, x = np.meshgrid(np.linspace(0, 1, my + 2 * mbc), np.linspace(0, 1, mx + 2 * mbc),
                       indexing="ij")
    q[..., 0] = 1.0 + np.exp(-40 * ((x - 0.4) ** 2 + (y - 0.55) ** 2))
    q[..., 1] = 0.2 * np.sin(3 * x + 1) * np.cos(2 * y)
    q[..., 2] = 0.1 * np.cos(5 * y) + 0.05 * rng.standard_normal(q.shape[:2])
    q[..., 0] += 0.1 * (x > 0.7)  # a jump so the limiter has work
    return q


def euler_problem(mx, my, seed):
    return P.euler_quadrants(mx, my, seed=seed)


def interior3(q, mbc=2):
    return q[mbc:-mbc, mbc:-mbc, mbc:-mbc]


def max_rel_err3(a, b, mbc=2):
    """max_rel_err for 3D arrays (z, y, x, m)."""
    a = interior3(a, mbc)
    b = interior3(b, mbc)
    top = max(np.abs(b[..., m]).max() for m in range(b.shape[-1]))
    worst = 0.0
    for m in range(a.shape[-1]):
        scale = max(np.abs(b[..., m]).max(), FIELD_FLOOR * top)
        d = np.abs(a[..., m] - b[..., m]).max()
        worst = max(worst, d / scale if scale > 0 else d)
    return worst


def random_layered(p, rng, maux=2):
    """Random q (ghost cells filled by the oracle's bc3) and random materials
    for an oracle3.Problem3."""
    from oracle import oracle3 as O3
    q = O3.bc3(p, rng.standard_normal(p.shape()))
    aux = np.empty(p.shape(maux))
    aux[..., 0] = rng.uniform(0.5, 3.0, aux.shape[:3])
    aux[..., 1] = rng.uniform(0.5, 2.0, aux.shape[:3])
    if maux > 2:
        aux[..., 2:] = 1.0
    return q, aux
