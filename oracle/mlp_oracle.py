"""CPU ORACLE of the learned limiter network -- test infrastructure only.

Float64 numpy restatement of the 5 -> 32 -> 32 -> 1 GELU network that the
learned limiter evaluates per interface wave (parameters as 1281 floats:
W1[5][32], b1[32], W2[32][32], b2[32], W3[32], b3), its TVD clip with MC
fallback, and the parameter gradients of the two training losses, written
out by hand (no autograd) so that the GPU forward and backward kernels can
be checked against it.  Pinned in tests by finite differences.

features() restates the feature pre-pass that feeds the network, on the C
oracle's Roe solver (oracle.rpn2): per interface wave the upwind ratio theta
of the reference's limiter (src/2d/inlinelimiter.f90:51-68: dotl/dotr over
wnorm2, upwind side chosen by the sign of s), the downwind ratio, the signed
Courant number, the wave strength relative to the two cell states and the
speed relative to the fastest wave of the interface.  Its theta is pinned
against the oracle's limiter (itself bit for bit with the reference) with
the Beam-Warming limiter, whose phi is theta; the other features against
closed forms (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

H = 32
SIZES = [("W1", (5, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(p, X):
    """X: (n, 5) features -> raw outputs (n,), and the activations."""
    P = unpack(p)
    z1 = X @ P["W1"] + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"] + P["b2"]
    h2 = gelu(z2)
    out = h2 @ P["W3"] + P["b3"][0]
    return out, (z1, h1, z2, h2)


def mc(theta):
    return np.maximum(0.0, np.minimum(np.minimum((1 + theta) / 2, 2.0), 2 * theta))


def superbee(theta):
    """max(0, min(1, 2 theta), min(2, theta)) (src/2d/inlinelimiter.f90:78-79)."""
    return np.maximum(0.0, np.maximum(np.minimum(1.0, 2 * theta), np.minimum(2.0, theta)))


def clip(raw, theta):
    """Sweby TVD region 0 <= phi <= min(2 theta, 2) (phi = 0 if theta <= 0);
    MC wherever the raw value falls outside.  Returns (phi, inside)."""
    hi = np.where(theta > 0, np.minimum(2 * theta, 2.0), 0.0)
    inside = (raw >= 0) & (raw <= hi) & (theta > 0)
    return np.where(inside, raw, mc(theta)), inside


def backward(p, X, dout):
    """Parameter gradient for upstream dL/dout (n,)."""
    P = unpack(p)
    out, (z1, h1, z2, h2) = forward(p, X)
    g = {"W3": h2.T @ dout, "b3": np.array([dout.sum()])}
    dz2 = np.outer(dout, P["W3"]) * gelu_d(z2)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(axis=0)
    dz1 = (dz2 @ P["W2"].T) * gelu_d(z1)
    g["W1"] = X.T @ dz1
    g["b1"] = dz1.sum(axis=0)
    return pack(g)


def grad_mc_regression(p, X, theta, target=mc):
    """L = sum (out - target(theta))^2 over the given waves (target: mc or
    superbee)."""
    out, _ = forward(p, X)
    e = out - target(theta)
    return backward(p, X, 2 * e), float((e * e).sum())


def grad_from_dphi(p, X, theta, dphi):
    """Backpropagate dL/dphi through the clip: no gradient where MC is used."""
    out, _ = forward(p, X)
    _, inside = clip(out, theta)
    return backward(p, X, np.where(inside, dphi, 0.0))


THETA_CLIP = 3.0


def _seqsum(terms):
    """Left-to-right sum starting from 0.0, as the kernels accumulate."""
    acc = np.zeros_like(terms[0])
    for t in terms:
        acc = acc + t
    return acc


def features(q, dx, dy, dt, solver_params, solver_id=2, mwaves=3, mbc=2):
    """Features of every interface wave of a 2D grid, in the learned
    limiter's table layout.

    q: (my+2mbc, mx+2mbc, meqn) with its ghost cells filled (reference
    layout).  Table slot (J+1)*(mx+4) + (i+1) for i in [-1, mx+2] and J in
    [-1, my+2]; direction 0 is the x-interface between cells (i-1, J) and
    (i, J) (live for 1 <= i <= mx+1, 0 <= J <= my+1), direction 1 the
    y-interface between (i, J-1) and (i, J) (live for 1 <= J <= my+1,
    0 <= i <= mx+1).  Returns feat (2, ntab, mwaves, 5) float32, theta
    (2, ntab, mwaves) float64 and nonzero (2, ntab, mwaves) bool; dead slots
    and zero waves hold zeros.

    For interface k with waves W_k and speeds s_k (rpn2 on the slice through
    the interface, as flux2 calls it):
      wnorm2 = W_k . W_k, dotl = W_{k-1} . W_k, dotr = W_k . W_{k+1}
      theta  = (s > 0 ? dotl : dotr) / wnorm2   (inlinelimiter.f90:51-68)
      f0 = clamp(theta, -3, 3) / 3, f1 = the same of the downwind ratio
           (s > 0 ? dotr : dotl) / wnorm2
      f2 = s dt/dx, f3 = |W_k| / (|q_{k-1}| + |q_k| + 1e-300),
      f4 = s / (max over waves |s| + 1e-300)
    with every sum accumulated left to right from 0 and the features
    rounded to float32 at the end.
    """
    from oracle import oracle as O

    q = np.ascontiguousarray(q, dtype=np.float64)
    M2, N2, meqn = q.shape
    mx, my = N2 - 2 * mbc, M2 - 2 * mbc
    ntab = (mx + 4) * (my + 4)
    feat = np.zeros((2, ntab, mwaves, 5), np.float32)
    theta = np.zeros((2, ntab, mwaves))
    nonzero = np.zeros((2, ntab, mwaves), bool)
    for d in (0, 1):
        dtd = dt / dx if d == 0 else dt / dy
        nslice = my + 2 if d == 0 else mx + 2  # slices 0 .. n+1
        for p in range(nslice):
            cells = q[p + mbc - 1] if d == 0 else q[:, p + mbc - 1]
            wave, s, _, _ = O.rpn2(solver_id, solver_params, d + 1, cells, mwaves, mbc=mbc)
            n = mx if d == 0 else my
            k = np.arange(1, n + 2)        # interfaces 1 .. n+1 along the slice
            e = k + mbc - 1                # their entries in the rpn2 arrays
            w0, w1, w2 = wave[e - 1], wave[e], wave[e + 1]
            sp = s[e]
            ql, qr = cells[e - 1], cells[e]
            smax = np.zeros(len(k))
            for mw in range(mwaves):
                smax = np.fmax(smax, np.abs(sp[:, mw]))
            nl = _seqsum([ql[:, m] * ql[:, m] for m in range(meqn)])
            nr = _seqsum([qr[:, m] * qr[:, m] for m in range(meqn)])
            qn = np.sqrt(nl) + np.sqrt(nr) + 1e-300
            if d == 0:
                slots = (p + 1) * (mx + 4) + (k + 1)
            else:
                slots = (k + 1) * (mx + 4) + (p + 1)
            for mw in range(mwaves):
                wn2 = _seqsum([w1[:, mw, m] * w1[:, mw, m] for m in range(meqn)])
                dl = _seqsum([w0[:, mw, m] * w1[:, mw, m] for m in range(meqn)])
                dr = _seqsum([w1[:, mw, m] * w2[:, mw, m] for m in range(meqn)])
                nz = wn2 != 0.0
                with np.errstate(divide="ignore", invalid="ignore"):
                    up = np.where(sp[:, mw] > 0.0, dl, dr) / wn2
                    dn = np.where(sp[:, mw] > 0.0, dr, dl) / wn2
                    f = [np.fmin(np.fmax(up, -THETA_CLIP), THETA_CLIP) / THETA_CLIP,
                         np.fmin(np.fmax(dn, -THETA_CLIP), THETA_CLIP) / THETA_CLIP,
                         sp[:, mw] * dtd,
                         np.sqrt(wn2) / qn,
                         sp[:, mw] / (smax + 1e-300)]
                for c in range(5):
                    feat[d, slots, mw, c] = np.where(nz, f[c], 0.0).astype(np.float32)
                theta[d, slots, mw] = np.where(nz, up, 0.0)
                nonzero[d, slots, mw] = nz
    return feat, theta, nonzero
