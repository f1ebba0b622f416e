"""CPU ORACLE of the learned limiter network -- test infrastructure only.

Float64 numpy restatement of the 5 -> 32 -> 32 -> 1 GELU network that the
learned limiter evaluates per interface wave (parameters as 1281 floats:
W1[5][32], b1[32], W2[32][32], b2[32], W3[32], b3), its TVD clip with MC
fallback, and the parameter gradients of the two training losses, written
out by hand (no autograd) so that the GPU forward and backward kernels can
be checked against it.  Pinned in tests by finite differences.
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
