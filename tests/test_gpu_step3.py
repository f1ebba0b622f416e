"""3D path on the GPU against the oracle (oracle/claw3d_oracle.c): bc3 and
the split step (dimsp3) bit for bit, the unsplit step (step3, every
method(3) family) to 1e-12 relative per field in fast arithmetic (the tile
and z-marching kernels sum each cell's increments in their own order) and
bit for bit in exact arithmetic and with a capacity function (the records
path), the claw3 loop (dt, Courant numbers, step counts exactly; q to
1e-12, or bit for bit with a capacity function), and the
heterogeneous-acoustics example's regression frame sums."""
import os

import numpy as np
import pytest

from classic_amd import problems as P
from classic_amd.gpu3 import Grid3D
from oracle import oracle3 as O3
from tests.helpers import cell_rel_errs, max_rel_err3, random_layered

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# sizes with ragged tiles in every direction (x strips of 60, 4x4 y-z tiles)
MX, MY, MZ = 70, 9, 6


def _grid(p, aux, q):
    g = Grid3D(p.mx, p.my, p.mz, maux=p.maux)
    g.set_aux(aux)
    g.set_q(q)
    return g


@pytest.mark.parametrize("mthbc", [(1, 1, 1, 1, 1, 1), (2, 2, 2, 2, 2, 2), (3, 3, 3, 3, 3, 3),
                                   (1, 3, 2, 2, 3, 0)])
def test_bc3_bitwise(mthbc):
    rng = np.random.default_rng(1)
    p = O3.Problem3(MX, MY, MZ, mthbc=mthbc, lower=(-0.3, -0.2, -1.0))
    q = rng.standard_normal(p.shape())
    aux = np.ones(p.shape(2))
    for t in (0.0, 0.02):
        ref = O3.bc3(p, q, t)
        with _grid(p, aux, q) as g:
            g.geom[:] = [p.geom.xlower, p.geom.ylower, p.geom.zlower, p.geom.dx, p.geom.dy, p.geom.dz]
            g.bc3(mthbc, t)
            np.testing.assert_array_equal(g.get_q(), ref)


@pytest.mark.parametrize("order,lim,fw", [(2, (3, 3), False), (2, (4, 1), False),
                                          (2, (2, 3), True), (1, (0, 0), False)])
def test_dimsp3_bitwise(order, lim, fw):
    rng = np.random.default_rng(2)
    p = O3.Problem3(MX, MY, MZ, mthlim=lim, method=(1, order, -1, 0, 0, 0, 0), use_fwave=fw)
    q, aux = random_layered(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=lim, use_fwave=fw)
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("m3,lim,fw", [(22, (3, 3), False), (22, (4, 2), False), (21, (1, 4), False),
                                       (20, (4, 4), False), (11, (0, 0), False), (10, (0, 0), False),
                                       (0, (3, 3), False), (22, (3, 3), True)])
def test_step3_matches_oracle(m3, lim, fw):
    rng = np.random.default_rng(3 + m3)
    order = 1 if lim == (0, 0) else 2
    p = O3.Problem3(MX, MY, MZ, mthlim=lim, method=(1, order, m3, 0, 0, 0, 0), use_fwave=fw)
    q, aux = random_layered(p, rng)
    dt = 0.25 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        cfl = g.step3(dt, method=p.method, mthlim=lim, use_fwave=fw)
        out = g.get_q()
    assert cfl == cfl_ref
    err = max_rel_err3(out, ref)
    print(f"step3 method {order},{m3} lim {lim} fwave {fw}: max rel err {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("table", [True, False])
def test_step3_every_tile_shape(tile, table):
    """Every tile shape and tile order of the unsplit kernel (with the
    material and coefficient tables where they are compiled, and reading
    aux) against the oracle, on a grid with ragged tiles in every
    direction."""
    rng = np.random.default_rng(40 + tile)
    p = O3.Problem3(MX, MY, MZ + 5, mthlim=(2, 2), method=(1, 2, 22, 0, 0, 0, 0))
    q, aux = _blocky_medium(p, rng, 3)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        g.set_tile(tile, (tile + int(table)) % 4)  # every tile order occurs
        g.set_aux_table(table)
        cfl = g.step3(dt, method=p.method, mthlim=(2, 2))
        out = g.get_q()
    assert cfl == cfl_ref
    err = max_rel_err3(out, ref)
    print(f"tile {tile} table {table}: max rel err {err:.2e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("tile", [5, 6])
@pytest.mark.parametrize("m3,lim,nmat", [(22, (2, 2), 2), (22, (3, 4), 7), (21, (1, 4), 3),
                                         (20, (4, 4), 2), (11, (0, 0), 3), (10, (0, 0), 5),
                                         (0, (3, 3), 2)])
def test_step3_zmarch_matches_oracle(tile, m3, lim, nmat):
    """The z-marching kernel (rows of 10 or 6 per block, 64 planes per
    block) against the oracle for every method(3) family, with the
    coefficient table (<= 4 materials) and without it: 13 rows (ragged
    row tiles), 70 planes (a second, ragged z segment) and 70 columns (a
    ragged second strip)."""
    rng = np.random.default_rng(60 + tile + m3)
    order = 1 if lim == (0, 0) else 2
    p = O3.Problem3(70, 13, 70, mthlim=lim, method=(1, order, m3, 0, 0, 0, 0),
                    mthbc=(3, 1, 2, 2, 3, 3))
    q, aux = _blocky_medium(p, rng, nmat)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    for order_bits in (0, 1):
        with _grid(p, aux, q) as g:
            g.set_tile(tile, order_bits)
            assert g.set_aux_table(True) == nmat
            cfl = g.step3(dt, method=p.method, mthlim=lim)
            out = g.get_q()
        assert cfl == cfl_ref
        err = max_rel_err3(out, ref)
        print(f"zmarch tile {tile} m3 {m3} nmat {nmat} order {order_bits}: max rel err {err:.2e}",
              flush=True)
        assert err <= TOL


@pytest.mark.parametrize("tile", [5, 6])
@pytest.mark.parametrize("m3,lim", [(22, (2, 2)), (21, (3, 4)), (11, (4, 4)), (20, (1, 1))])
def test_step3_zmarch_layered_medium(tile, m3, lim):
    """Materials that vary along z only (layers of 1 to 5 planes, three
    materials): most waves see one material in every 3x3 neighbourhood and
    take the kernel's one-material form of the transverse records, the
    waves at a layer interface the general form."""
    rng = np.random.default_rng(80 + tile + m3)
    p = O3.Problem3(70, 13, 70, mthlim=lim, method=(1, 2, m3, 0, 0, 0, 0),
                    mthbc=(3, 3, 3, 3, 3, 3))
    q = O3.bc3(p, rng.standard_normal(p.shape()))
    pairs = np.array([[1.0, 1.0], [6.0, 2.0], [2.5, 0.7]])
    nz = p.shape(2)[0]
    layer = np.repeat(np.arange(nz), rng.integers(1, 6, nz))[:nz] % 3
    aux = np.ascontiguousarray(np.broadcast_to(pairs[layer][:, None, None, :], p.shape(2)))
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        g.set_tile(tile, 1)
        assert g.set_aux_table(True) == 3
        cfl = g.step3(dt, method=p.method, mthlim=lim)
        out = g.get_q()
    assert cfl == cfl_ref
    err = max_rel_err3(out, ref)
    print(f"zmarch layered tile {tile} m3 {m3}: max rel err {err:.2e}", flush=True)
    assert err <= TOL


def test_step3_larger_grid_many_tiles():
    rng = np.random.default_rng(9)
    p = O3.Problem3(125, 34, 23, mthlim=(3, 4))
    q, aux = random_layered(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, _ = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        g.step3(dt, method=p.method, mthlim=(3, 4))
        out = g.get_q()
    assert max_rel_err3(out, ref) <= TOL


def test_step3_benchmark_size_windows_match_oracle():
    """At the benchmark size (512^3, the two-layer medium, walls, unsplit
    method 22, superbee), one GPU step on random data agrees with the oracle
    run on 24^3 windows of the same state: a cell's update only reads cells
    within two of it, so a window with two ghost layers taken from the full
    state reproduces the full step on its interior.  Windows: an interior
    one, one across the material interface at z = 0, the corner at three
    walls and the opposite corner (the ragged last strip and tiles)."""
    n, w, mbc = 512, 24, 2
    rng = np.random.default_rng(512)
    full = O3.Problem3(n, n, n, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2),
                       method=(1, 2, 22, 0, 0, 0, 0))
    q = O3.bc3(full, rng.standard_normal(full.shape()))
    aux = P.layered_medium_aux(n, n, n, mbc=mbc, below=(1.0, 1.0), above=(6.0, 2.0))
    dx = 2.0 / n
    dt = 0.4 * dx / 2.0
    with Grid3D(n, n, n) as g:
        g.set_aux(aux)
        g.set_q(q)
        g.step3(dt, method=(1, 2, 22, 0, 0, 0, 0), mthlim=(2, 2))
        out = g.get_q()
    print("512^3 step done", flush=True)
    corners = [(201, 137, 310), (250, 300, n // 2 - w // 2 + 1), (1, 1, 1),
               (n - w + 1, n - w + 1, n - w + 1)]
    for i0, j0, k0 in corners:
        # Fortran cell k lives at array index k + mbc - 1
        sl = tuple(slice(c - 1, c + w + 3) for c in (k0, j0, i0))
        lower = tuple(-1.0 + (c - 1) * dx for c in (i0, j0, k0))
        upper = tuple(lo + w * dx for lo in lower)
        win = O3.Problem3(w, w, w, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2),
                          method=(1, 2, 22, 0, 0, 0, 0), lower=lower, upper=upper)
        ref, _ = O3.step3(win, q[sl], aux[sl], dt)
        err = max_rel_err3(out[sl], ref)
        print(f"window at ({i0}, {j0}, {k0}): max rel err {err:.2e}", flush=True)
        assert np.abs(ref[2:-2, 2:-2, 2:-2] - q[sl][2:-2, 2:-2, 2:-2]).max() > 1e-3
        assert err <= TOL


def test_heterogeneous_example_frames_and_regression_sums():
    gold = np.load(os.path.join(GOLD, "hetero3d_frames_20.npz"))
    p = O3.Problem3(20, 20, 20, mthbc=(1, 1, 1, 1, 1, 0), mthlim=(3, 3))
    q, aux = P.acoustics3d_heterogeneous(20)
    ref = O3.frames(p, q, aux, 0.0, 1.0, 2, 0.01)
    with _grid(p, aux, q) as g:
        out = g.frames(0.0, 1.0, 2, 0.01, mthlim=(3, 3), mthbc=(1, 1, 1, 1, 1, 0))
    for k, ((t_r, q_r), (t_g, q_g)) in enumerate(zip(ref, out)):
        assert t_r == t_g
        err = max_rel_err3(q_g, q_r)
        sums = np.array([q_g[2:-2, 2:-2, 2:-2, m].sum() for m in range(3)])
        reg = gold["regression_sums"][k]
        print(f"frame {k + 1}: t={t_g} max rel err {err:.2e} sums {sums} regression {reg}")
        assert err <= TOL
        assert abs(sums[0] - reg[0]) <= 1e-13 * abs(reg[0])
        assert np.all(np.abs(sums[1:] - reg[1:]) <= 1e-14)


@pytest.mark.parametrize("m3", [22, -1])
def test_claw3_run_with_retakes_matches_oracle(m3):
    p = O3.Problem3(36, 28, 20, mthbc=(3, 3, 3, 3, 3, 3), mthlim=(2, 2),
                    method=(1, 2, m3, 0, 0, 0, 0))
    qc, _, _ = P.acoustics3d_pulse(36, seed=4, width=0.3)
    q = qc[: p.shape()[0], : p.shape()[1]].copy()
    aux = P.layered_medium_aux(36, 28, 20, below=(1.0, 1.0), above=(6.0, 2.0))
    r = O3.claw3(p, q, aux, 0.0, 0.4, [0.2, 1e9], [1.0, 0.9], 500)
    with _grid(p, aux, q) as g:
        rg = g.claw3(0.0, 0.4, [0.2, 1e9], [1.0, 0.9], 500, method=p.method, mthlim=(2, 2),
                     mthbc=(3, 3, 3, 3, 3, 3))
        out = g.get_q()
    assert rg.info == r.info == 0
    assert rg.tend == r.tend
    np.testing.assert_array_equal(rg.nv, r.nv)
    np.testing.assert_array_equal(rg.dtv, r.dtv)
    np.testing.assert_array_equal(rg.cflv, r.cflv)
    err = max_rel_err3(out, r.q)
    print(f"claw3 method(3)={m3}: {r.nv[1]} steps, max rel err {err:.2e}")
    if m3 < 0:
        np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], r.q[2:-2, 2:-2, 2:-2])
    assert err <= TOL


def test_unsupported_options_fail_loudly():
    from classic_amd._lib import ClawGpuError
    with Grid3D(8, 8, 8) as g:
        with pytest.raises(ClawGpuError):
            g.step3(0.01, method=(1, 2, 22, 0, 0, 3, 2))   # capacity index beyond maux
        with pytest.raises(ClawGpuError):
            g.step3(0.01, method=(1, 2, 12, 0, 0, 0, 2))   # not a chkmth method
        with pytest.raises(ClawGpuError):
            g.bc3((0, 1, 1, 1, 1, 1))                       # user BC off z-hi


def _blocky_medium(p, rng, nmat):
    """q random, materials drawn from nmat random (Z, c) pairs assigned in
    blocks of 3x2x2 cells (ghost cells included)."""
    q = O3.bc3(p, rng.standard_normal(p.shape()))
    pairs = np.stack([rng.uniform(0.5, 3.0, nmat), rng.uniform(0.5, 2.0, nmat)], axis=1)
    nz, ny, nx, _ = p.shape(2)
    kk, jj, ii = np.meshgrid(np.arange(nz) // 2, np.arange(ny) // 2, np.arange(nx) // 3,
                             indexing="ij")
    which = (kk * 7 + jj * 3 + ii * 5 + rng.integers(0, nmat, (nz, ny, nx))) % nmat
    aux = pairs[which]
    return q, np.ascontiguousarray(aux)


@pytest.mark.parametrize("nmat", [3, 7])
@pytest.mark.parametrize("m3,lim", [(22, (2, 2)), (21, (3, 4)), (10, (1, 1))])
def test_material_table_bitwise_with_aux_path(m3, lim, nmat):
    """The one-byte material index reads the same doubles as aux, and with
    at most 4 materials the tabulated transverse coefficients are the bits
    kcoef computes: the unsplit tile kernel is bit for bit the same with the
    tables on and off, and within 1e-12 of the oracle.  (The z-marching
    kernel, the default, reads the table only: without it the tile kernel
    runs.)"""
    rng = np.random.default_rng(40 + m3 + nmat)
    p = O3.Problem3(MX, MY, MZ, mthlim=lim, method=(1, 2, m3, 0, 0, 0, 0))
    q, aux = _blocky_medium(p, rng, nmat)
    dt = 0.25 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    outs = []
    for table in (True, False):
        with _grid(p, aux, q) as g:
            g.set_tile(4, 2)
            assert g.set_aux_table(table) == nmat
            cfl = g.step3(dt, method=p.method, mthlim=lim)
            outs.append(g.get_q())
        assert cfl == cfl_ref
    np.testing.assert_array_equal(outs[0], outs[1])
    assert max_rel_err3(outs[0], ref) <= TOL


def test_material_table_limit():
    """64 distinct (Z, c) pairs use the table; 65 fall back to aux."""
    p = O3.Problem3(20, 6, 5)
    for nmat, want in ((64, 64), (65, 0)):
        q, aux = _blocky_medium(p, np.random.default_rng(nmat), nmat)
        assert len({(a, b) for a, b in aux.reshape(-1, 2)}) == nmat
        with _grid(p, aux, q) as g:
            assert g.set_aux_table(True) == want


def test_config3_full_size_100_steps_symmetry_and_last_step():
    """Config 3 at its size and length: 512^3 layered acoustics, walls,
    superbee, unsplit method 22, CFL 0.9, 100 accepted steps of the claw3
    loop, with the pulse centred in x and y (at z = 0.3, above the
    interface) so that the medium and the data are mirror-symmetric in x and
    y.  (i) The run takes its 100 steps; (ii) after them the solution is
    mirror-symmetric in x (p, v, w even, u odd) and in y (p, u, w even, v
    odd) to 1e-12 of each field's largest value -- an index or tile-edge
    slip anywhere breaks it; (iii) the 100th step agrees with the oracle run
    on 24^3 windows of the 99-step state (on the wave fronts and across the
    material interface)."""
    n, w, mbc = 512, 24, 2
    q, aux, _ = P.acoustics3d_pulse(n, center=(0.0, 0.0, 0.3), width=0.05)
    dx = 2.0 / n
    method, mthlim, mthbc = (1, 2, 22, 0, 0, 0, 2), (2, 2), (3, 3, 3, 3, 3, 3)
    with Grid3D(n, n, n) as g:
        g.set_aux(aux)
        g.set_q(q)
        del q
        r = g.claw3(0.0, 1e30, [0.9 * dx / 2.0, 1e99], [1.0, 0.9], 99, method=method,
                    mthlim=mthlim, mthbc=mthbc)
        assert r.info == 11 and r.nv[1] == 99  # stopped by the step count, as asked
        q99 = g.get_q()
        r2 = g.claw3(r.tend, 1e30, [r.dtv[4], 1e99], [1.0, 0.9], 1, method=method,
                     mthlim=mthlim, mthbc=mthbc)
        assert r2.nv[1] == 1
        dt = r2.dtv[2]
        q100 = g.get_q()
    print(f"512^3: 100 steps to t = {r2.tend:.4f}", flush=True)
    inner = q100[2:-2, 2:-2, 2:-2]
    for axis, odd in ((2, 1), (1, 2)):  # array axes (z, y, x, m): x is 2, y is 1
        flip = np.flip(inner, axis=axis)
        for m in range(4):
            top = np.abs(inner[..., m]).max()
            sgn = -1.0 if m == odd else 1.0
            err = np.abs(inner[..., m] - sgn * flip[..., m]).max() / top
            print(f"mirror axis {axis} field {m}: {err:.2e} of {top:.3e}", flush=True)
            assert err <= TOL
        del flip
    # windows of the last step: where the pressure and where the vertical
    # velocity peak, the centre column across the interface at z = 0, and the
    # front where it crosses the interface off-centre
    def around(field):
        kz, jy, ix = np.unravel_index(np.argmax(np.abs(inner[..., field])), inner.shape[:3])
        return tuple(min(max(int(c) + 1 - w // 2, 1), n - w + 1) for c in (ix, jy, kz))

    h = n // 2 - w // 2 + 1
    corners = [around(0), around(3), (h, h, h), (h + 40, h, h)]
    full = O3.Problem3(n, n, n, mthbc=mthbc, mthlim=mthlim, method=method)
    q99 = O3.bc3(full, q99)
    for i0, j0, k0 in corners:
        sl = tuple(slice(c - 1, c + w + 3) for c in (k0, j0, i0))
        lower = tuple(-1.0 + (c - 1) * dx for c in (i0, j0, k0))
        upper = tuple(lo + w * dx for lo in lower)
        win = O3.Problem3(w, w, w, mthbc=mthbc, mthlim=mthlim, method=method, lower=lower,
                          upper=upper)
        ref, _ = O3.step3(win, q99[sl], aux[sl], dt)
        err = max_rel_err3(q100[sl], ref)
        print(f"window at ({i0}, {j0}, {k0}): max rel err {err:.2e}", flush=True)
        assert np.abs(ref[2:-2, 2:-2, 2:-2] - q99[sl][2:-2, 2:-2, 2:-2]).max() > 1e-8
        assert err <= TOL


def test_config3_full_size_zmarch_vs_tile_kernel_100_steps():
    """Config 3 at its size and length run by the two unsplit kernels, each
    held to the oracle on small grids above: the z-marching kernel (the
    default) and the 60x2x4 tile kernel, which share no code beyond the
    interface solves and sum the increments in other orders.  Both take the
    same 100 steps to the same dt; the states agree to 1e-12 per field."""
    n = 512
    q, aux, _ = P.acoustics3d_pulse(n, seed=0)
    dx = 2.0 / n
    method, mthlim, mthbc = (1, 2, 22, 0, 0, 0, 2), (2, 2), (3, 3, 3, 3, 3, 3)
    res = []
    for tile, order in ((6, 3), (4, 2)):
        with Grid3D(n, n, n) as g:
            g.set_tile(tile, order)
            g.set_aux(aux)
            g.set_q(q)
            r = g.claw3(0.0, 1e30, [0.9 * dx / 2.0, 1e99], [1.0, 0.9], 100, method=method,
                        mthlim=mthlim, mthbc=mthbc)
            assert r.nv[1] == 100
            res.append((r, g.get_q()))
        print(f"tile {tile}: 100 steps to t = {r.tend:.6f}", flush=True)
    del q, aux
    (ra, qa), (rb, qb) = res
    dterr = abs(ra.dtv[4] - rb.dtv[4]) / rb.dtv[4]
    err = max_rel_err3(qa, qb)
    cerr = max(cell_rel_errs(qa, qb))
    print(f"z-marching vs tile kernel 512^3 x 100: max rel err {err:.2e}, cell rel err "
          f"{cerr:.2e}, dt rel err {dterr:.2e}", flush=True)
    assert err <= TOL and dterr <= TOL


# ------------------------------------------- exact mode and capacity function

def _gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("case", range(10))
def test_exact_and_capacity_steps_equal_the_reference_build(case):
    """The golden single steps written by the reference Fortran
    (tests/golden/make_golden3.py: every method(3) family, f-waves, the
    split step, and the capacity function in aux(3) for the unsplit step
    (case 6) and the split step (case 9)): the exact unsplit step (the
    records path) and the split step equal the reference's q bit for bit,
    ghost cells excluded for the unsplit step (the reference leaves partial
    updates there), and the Courant number exactly."""
    d = _gold(f"step3_case{case}")
    method = tuple(int(x) for x in d["method"])
    lim = tuple(int(x) for x in d["mthlim"])
    fw = bool(int(d["use_fwave"]))
    maux = d["aux"].shape[-1]
    p = O3.Problem3(6, 5, 4, maux=maux, mthlim=lim, method=method, use_fwave=fw)
    dt = 0.02
    with _grid(p, d["aux"], d["q_in"]) as g:
        g.set_arith(0)
        if method[2] >= 0:
            cfl = g.step3(dt, method=method, mthlim=lim, use_fwave=fw)
        else:
            cfl = g.dimsp3(dt, method=method, mthlim=lim, use_fwave=fw)
        out = g.get_q()
    print(f"case {case}: method {method[:3]} capa {method[5]} cfl {cfl!r}", flush=True)
    assert cfl == float(d["cfl"])
    if method[2] >= 0:
        np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], d["q_out"][2:-2, 2:-2, 2:-2])
    else:
        np.testing.assert_array_equal(out, d["q_out"])


def _capa_state(p, rng):
    q, aux2 = random_layered(p, rng)
    aux = np.empty(p.shape(3))
    aux[..., :2] = aux2
    aux[..., 2] = rng.uniform(0.5, 1.5, aux.shape[:3])
    return q, aux


@pytest.mark.parametrize("m3,lim,fw", [(22, (3, 3), False), (22, (4, 2), False), (21, (1, 4), False),
                                       (20, (4, 4), False), (11, (0, 0), False), (10, (0, 0), False),
                                       (0, (3, 3), False), (22, (3, 3), True)])
def test_step3_capacity_bitwise_vs_oracle(m3, lim, fw):
    """Unsplit step with the capacity function in aux(3) on a grid of
    ragged sides (70 x 9 x 6; blocks of slices with odd counts) against the
    oracle (pinned to the reference build by the golden vectors): q and the
    Courant number bit for bit, every method(3) family."""
    rng = np.random.default_rng(80 + m3)
    order = 1 if lim == (0, 0) else 2
    p = O3.Problem3(MX, MY, MZ, maux=3, mthlim=lim, method=(1, order, m3, 0, 0, 3, 0),
                    use_fwave=fw)
    q, aux = _capa_state(p, rng)
    dt = 0.25 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        cfl = g.step3(dt, method=p.method, mthlim=lim, use_fwave=fw)
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], ref[2:-2, 2:-2, 2:-2])


@pytest.mark.parametrize("order,lim", [(2, (3, 3)), (2, (4, 1)), (1, (0, 0))])
def test_dimsp3_capacity_bitwise(order, lim):
    """Split step with the capacity function: bit for bit with the oracle,
    ghost cells included, and the Courant number."""
    rng = np.random.default_rng(90 + order)
    p = O3.Problem3(MX, MY, MZ, maux=3, mthlim=lim, method=(1, order, -1, 0, 0, 3, 0))
    q, aux = _capa_state(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.dimsp3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        cfl = g.dimsp3(dt, method=p.method, mthlim=lim)
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out, ref)


def test_exact_mode_equals_oracle_without_capacity():
    """Without a capacity function the exact mode (set_arith(0)) is the same
    records path: bit for bit with the oracle on a layered medium, where the
    fast kernels agree to 1e-12 only."""
    rng = np.random.default_rng(99)
    p = O3.Problem3(MX, MY + 4, MZ + 3, mthlim=(2, 2), method=(1, 2, 22, 0, 0, 0, 0))
    q, aux = random_layered(p, rng)
    dt = 0.3 * p.geom.dx / 2.0
    ref, cfl_ref = O3.step3(p, q, aux, dt)
    with _grid(p, aux, q) as g:
        g.set_arith(0)
        cfl = g.step3(dt, method=p.method, mthlim=(2, 2))
        out = g.get_q()
    assert cfl == cfl_ref
    np.testing.assert_array_equal(out[2:-2, 2:-2, 2:-2], ref[2:-2, 2:-2, 2:-2])


@pytest.mark.parametrize("split", [False, True])
def test_claw3_with_capacity_matches_oracle(split):
    """The claw3 loop with a capacity function (variable time steps from its
    Courant numbers, retakes on a too-large first dt) against the oracle's
    claw3: dt, Courant numbers and step counts exactly, q bit for bit."""
    rng = np.random.default_rng(100 + int(split))
    m3 = -1 if split else 22
    p = O3.Problem3(24, 14, 10, maux=3, mthlim=(3, 4), method=(1, 2, m3, 0, 0, 3, 0),
                    mthbc=(1, 3, 2, 2, 3, 1))
    q, aux = _capa_state(p, rng)
    dtv, cflv = (0.5, 1.0), (1.0, 0.9)
    r = O3.claw3(p, q, aux, 0.0, 0.3, dtv, cflv, 40)
    with _grid(p, aux, q) as g:
        out = g.claw3(0.0, 0.3, dtv, cflv, 40, method=p.method, mthlim=(3, 4), mthbc=p.mthbc)
        qg = g.get_q()
    print(f"split {split}: steps {out.nv[1]} info {out.info} dtv {out.dtv}", flush=True)
    assert out.info == r.info and out.nv[1] == r.nv[1] and out.nv[1] > 3
    np.testing.assert_array_equal(out.dtv, r.dtv)
    np.testing.assert_array_equal(out.cflv, r.cflv)
    np.testing.assert_array_equal(qg[2:-2, 2:-2, 2:-2], r.q[2:-2, 2:-2, 2:-2])
