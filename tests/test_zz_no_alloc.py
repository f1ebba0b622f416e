"""clawgpu_step2_noalloc, the context-free step for callers that cannot
allocate device memory, against the existing entry point clawgpu_step2 (a
context, the same fused kernel): two calls in a row through the C-ABI, the
second on a smaller grid with a smaller Courant number, so a Courant
accumulator left over from the first call would show in the second's cfl.
EXACT arithmetic is held bit for bit (q interior and cfl), FAST to the same
kernel's results bit for bit as well (same instructions per cell; only the
row pitch and plane size differ).  The caller's buffers use a pitch and plane
size other than the context's, and qnew's ghost cells must come back as the
caller left them."""
import ctypes as C

import numpy as np
import pytest

from classic_amd import SOLVER_ACOUSTICS, SOLVER_EULER, gpu
from classic_amd import problems as P
from classic_amd._lib import ERR_ARG, ERR_UNSUPPORTED, load
from oracle import oracle as O
from tests.helpers import smooth_acoustics

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

METHOD = (1, 2, 2, 0, 0, 0, 0)
SENTINEL = -7.25


def _case(solver, mx, my, mbc):
    if solver == SOLVER_EULER:
        q, dx, dy = P.euler_quadrants(mx, my, seed=11, mbc=mbc)
        p = O.Problem(4, 3, mx, my, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                      method=METHOD, mthlim=[4, 4, 4], mthbc=(1, 3, 2, 1), mbc=mbc)
        return O.bc2(p, q), dx, dy, [1.4, 1.0], [4, 4, 4]
    q = smooth_acoustics(mx, my, seed=3, mbc=mbc)
    p = O.Problem(3, 2, mx, my, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                  method=METHOD, mthlim=[4, 2], mthbc=(3, 3, 2, 2), mbc=mbc)
    return O.bc2(p, q), 1.0 / mx, 1.0 / my, [1.0, 4.0], [4, 2]


def _to_planes(q, pitch, plane):
    """reference layout [j][i][m] -> field planes m*plane + j*pitch + i"""
    ny, nx, meqn = q.shape
    out = np.full(meqn * plane, SENTINEL)
    for m in range(meqn):
        v = out[m * plane: m * plane + ny * pitch].reshape(ny, pitch)
        v[:, :nx] = q[..., m]
    return out


def _from_planes(buf, shape, pitch, plane):
    ny, nx, meqn = shape
    q = np.empty(shape)
    for m in range(meqn):
        q[..., m] = buf[m * plane: m * plane + ny * pitch].reshape(ny, pitch)[:, :nx]
    return q


def _noalloc(lib, solver, params, arith, q, mbc, dx, dy, dt, mthlim, pad):
    ny, nx, meqn = q.shape
    mx, my = nx - 2 * mbc, ny - 2 * mbc
    pitch = nx + pad
    plane = pitch * ny + 3 * pad
    dq = torch.from_numpy(_to_planes(q, pitch, plane)).cuda()
    dn = torch.full_like(dq, SENTINEL)
    dcfl = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    par = np.ascontiguousarray(params, dtype=np.float64)
    meth = np.ascontiguousarray(METHOD, dtype=np.int32)
    lim = np.ascontiguousarray(mthlim, dtype=np.int32)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.clawgpu_step2_noalloc(
        solver, par.ctypes.data_as(C.POINTER(C.c_double)), par.size, arith, mx, my, mbc, pitch,
        plane, dq.data_ptr(), dn.data_ptr(), dx, dy, dt, meth.ctypes.data_as(C.POINTER(C.c_int)),
        lim.ctypes.data_as(C.POINTER(C.c_int)), 0, dcfl.data_ptr(), stream)
    assert rc == 0, lib.clawgpu_last_error().decode()
    torch.cuda.synchronize()
    buf = dn.cpu().numpy()
    qn = _from_planes(buf, q.shape, pitch, plane)
    # ghost cells and padding of qnew are the caller's: untouched
    inside = np.zeros(q.shape, dtype=bool)
    inside[mbc:-mbc, mbc:-mbc] = True
    assert np.all(qn[~inside] == SENTINEL)
    mask = np.ones(buf.size, dtype=bool)
    for m in range(meqn):
        mask[m * plane: m * plane + ny * pitch].reshape(ny, pitch)[:, :nx] = False
    assert np.all(buf[mask] == SENTINEL)
    return qn, float(dcfl.cpu()[0])


def _with_context(solver, params, arith, q, mbc, dx, dy, dt, mthlim):
    ny, nx, _ = q.shape
    with gpu.Grid2D(solver, nx - 2 * mbc, ny - 2 * mbc, params, mbc=mbc, arith=arith) as g:
        g.set_q(q)
        cfl = g.step2(dx, dy, dt, METHOD, mthlim)
        return g.get_q(), cfl


@pytest.mark.parametrize("arith", [0, 1])
@pytest.mark.parametrize("solver", [SOLVER_EULER, SOLVER_ACOUSTICS])
def test_noalloc_matches_step2_two_calls(solver, arith):
    lib = load()
    runs = [  # (mx, my, mbc, dt factor, pad): the second grid smaller, its cfl smaller
        (187, 131, 2, 0.8, 5),
        (61, 23, 3, 0.3, 1),
    ]
    cfls = []
    for mx, my, mbc, f, pad in runs:
        q, dx, dy, params, mthlim = _case(solver, mx, my, mbc)
        dt = f * dx / (3.0 if solver == SOLVER_EULER else 2.0)
        qn, cfl = _noalloc(lib, solver, params, arith, q, mbc, dx, dy, dt, mthlim, pad)
        qc, cflc = _with_context(solver, params, arith, q, mbc, dx, dy, dt, mthlim)
        i = (slice(mbc, -mbc), slice(mbc, -mbc))
        assert np.array_equal(qn[i], qc[i]), f"{mx}x{my}: q differs from clawgpu_step2"
        assert cfl == cflc, (cfl, cflc)
        cfls.append(cfl)
    assert 0.0 < cfls[1] < cfls[0]


def test_noalloc_rejections():
    """Argument errors come back as error codes before anything is enqueued."""
    lib = load()
    dq = torch.zeros(4 * 64 * 64, dtype=torch.float64, device="cuda")
    dq[: 64 * 64] = 1.0  # density 1 at rest, energy 2.5: a valid uniform state
    dq[3 * 64 * 64:] = 2.5
    dn = torch.zeros_like(dq)
    dc = torch.zeros(1, dtype=torch.float64, device="cuda")
    par = np.array([1.4, 1.0])
    meth = np.ascontiguousarray(METHOD, dtype=np.int32)

    def call(mx=20, my=20, mbc=2, pitch=64, plane=64 * 64, qn=None, lim=(4, 4, 4), solver=SOLVER_EULER):
        lim_a = np.ascontiguousarray(lim, dtype=np.int32)
        return lib.clawgpu_step2_noalloc(
            solver, par.ctypes.data_as(C.POINTER(C.c_double)), 2, 0, mx, my, mbc, pitch, plane,
            dq.data_ptr(), (qn if qn is not None else dn).data_ptr(), 0.01, 0.01, 0.001,
            meth.ctypes.data_as(C.POINTER(C.c_int)), lim_a.ctypes.data_as(C.POINTER(C.c_int)), 0,
            dc.data_ptr(), None)

    assert call(mx=(1 << 20) + 1, pitch=(1 << 21), plane=(1 << 22)) == ERR_ARG
    assert call(pitch=23) == ERR_ARG
    assert call(plane=64 * 23) == ERR_ARG
    assert call(qn=dq) == ERR_ARG
    assert call(mbc=1) == ERR_ARG
    assert call(lim=(6, 6, 6)) == ERR_UNSUPPORTED
    assert call(solver=5) == ERR_UNSUPPORTED
    assert call() == 0
    torch.cuda.synchronize()
    # at rest with p = 0.4 * 2.5 = 1: the outer speeds are -+c, c^2 = gamma p / rho = 1.4
    assert abs(dc.item() - 0.1 * np.sqrt(1.4)) <= 1e-15
