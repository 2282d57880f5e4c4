"""Oracle: grid sound speed and the Gaussian-beam ODEs in a heterogeneous
medium (test infrastructure only).

Restates ``beamax/utils/interp.py:101-270`` (tensor-product n-linear and
cubic B-spline interpolants with SciPy's spline prefilter, boundary index
maps) with closed-form derivatives of the weights, and
``beamax/gb/gb_solvers.py:787-908`` (``riccati_rhs``,
``coupled_rhs_absorption``), integrated with SciPy's DOP853 at tight
tolerances.
"""

from __future__ import annotations

import itertools

import numpy as np


def prepare(c_map, method="linear", boundary="nearest", smooth_sigma=None):
    """interp.py:31-75: optional Gaussian smoothing, B-spline prefilter."""
    from scipy import ndimage
    v = np.asarray(c_map, dtype=float)
    if smooth_sigma is not None:
        v = ndimage.gaussian_filter(v, sigma=smooth_sigma, mode=boundary)
    if method == "bspline3":
        return ndimage.spline_filter(v, order=3, mode=boundary)
    return v


def map_index(i, n, boundary):
    """interp.py:78-98."""
    i = np.asarray(i)
    if n == 1:
        return np.zeros_like(i)
    if boundary == "nearest":
        return np.clip(i, 0, n - 1)
    if boundary == "wrap":
        return np.mod(i, n)
    if boundary == "reflect":
        r = np.mod(i, 2 * n)
        return np.where(r < n, r, 2 * n - 1 - r)
    if boundary == "mirror":
        r = np.mod(i, 2 * n - 2)
        return np.where(r < n, r, 2 * n - 2 - r)
    raise ValueError(boundary)


def _weights(t, method):
    """weights over the offsets and their first/second t-derivatives."""
    if method == "linear":
        w = np.stack([1 - t, t], -1)
        w1 = np.stack([-np.ones_like(t), np.ones_like(t)], -1)
        w2 = np.zeros_like(w)
        return (0, 1), w, w1, w2
    s = 1 - t
    w = np.stack([s ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6,
                  (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6], -1)
    w1 = np.stack([-s ** 2 / 2, (9 * t ** 2 - 12 * t) / 6, (-9 * t ** 2 + 6 * t + 3) / 6,
                   t ** 2 / 2], -1)
    w2 = np.stack([s, 3 * t - 2, -3 * t + 1, t], -1)
    return (-1, 0, 1, 2), w, w1, w2


def interp_eval(table, spacing, origin, method, boundary, x):
    """c, grad (n, d) and Hessian (n, d, d) at points x (n, d)."""
    table = np.asarray(table, dtype=float)
    x = np.atleast_2d(np.asarray(x, dtype=float))
    n, d = x.shape
    u = (x - np.asarray(origin)) / np.asarray(spacing)
    i0 = np.floor(u).astype(np.int64)
    t = u - i0
    offs, w, w1, w2 = _weights(t, method)
    c = np.zeros(n)
    g = np.zeros((n, d))
    H = np.zeros((n, d, d))
    for combo in itertools.product(range(len(offs)), repeat=d):
        idx = tuple(map_index(i0[:, a] + offs[combo[a]], table.shape[a], boundary)
                    for a in range(d))
        v = table[idx]
        W = [w[:, a, combo[a]] for a in range(d)]
        W1 = [w1[:, a, combo[a]] for a in range(d)]
        W2 = [w2[:, a, combo[a]] for a in range(d)]
        c += np.prod(W, axis=0) * v
        for a in range(d):
            pa = W1[a] * np.prod([W[b] for b in range(d) if b != a], axis=0)
            g[:, a] += pa * v
            for b in range(d):
                if a == b:
                    pab = W2[a] * np.prod([W[e] for e in range(d) if e != a], axis=0)
                else:
                    pab = W1[a] * W1[b] * np.prod([W[e] for e in range(d) if e not in (a, b)],
                                                  axis=0)
                H[:, a, b] += pab * v
    sp = np.asarray(spacing)
    return c, g / sp, H / (sp[:, None] * sp[None, :])


class Speed:
    def __init__(self, c_map, spacing, origin=None, method="linear", boundary="nearest",
                 smooth_sigma=None):
        self.table = prepare(c_map, method, boundary, smooth_sigma)
        self.spacing = np.asarray(spacing, float)
        self.origin = np.zeros(self.table.ndim) if origin is None else np.asarray(origin, float)
        self.method, self.boundary = method, boundary

    def __call__(self, x):
        return interp_eval(self.table, self.spacing, self.origin, self.method, self.boundary, x)


def rhs(y, mode, speed, d, lam=0.0):
    """coupled_rhs_absorption (gb_solvers.py:820-867) with riccati_rhs
    (787-817) on the real state (x, p, Re/Im M, Re/Im A)."""
    x = y[:d]
    p = y[d:2 * d]
    M = y[2 * d:2 * d + d * d].reshape(d, d) + 1j * y[2 * d + d * d:2 * d + 2 * d * d].reshape(d, d)
    A = y[-2] + 1j * y[-1]
    c, g, H = speed(x[None])
    c, g, H = c[0], g[0], H[0]
    npn = np.linalg.norm(p)
    G = mode * c * npn
    Gx = mode * g * npn
    Gp = mode * c * p / npn
    Gpp = mode * c * (np.eye(d) / npn - np.outer(p, p) / npn ** 3)
    Gxp = mode * np.outer(g, p) / npn
    Gxx = mode * H * npn
    dM = -(Gxx + Gxp @ M + M @ Gxp.T + M @ Gpp @ M)
    dA = -A * (c ** 2 * np.trace(M) - Gx @ Gp - Gp @ M @ Gp + lam * G) / (2 * G)
    return np.concatenate([Gp, -Gx, dM.real.ravel(), dM.imag.ravel(), [dA.real, dA.imag]])


def integrate(x0, p0, M0, a0, mode, ts, speed, lam=0.0, rtol=1e-11, atol=1e-13):
    """One beam from ts[0], saved at ts: (xt, pt, Mt, At) with shapes
    (Nt, d), (Nt, d), (Nt, d, d), (Nt,)."""
    from scipy.integrate import solve_ivp
    d = len(x0)
    M0 = np.asarray(M0, complex)
    y0 = np.concatenate([x0, p0, M0.real.ravel(), M0.imag.ravel(), [a0.real, a0.imag]])
    ts = np.asarray(ts, float)
    sol = solve_ivp(lambda t, y: rhs(y, mode, speed, d, lam), (ts[0], ts[-1]), y0,
                    method="DOP853", t_eval=ts, rtol=rtol, atol=atol)
    Y = sol.y.T
    xt = Y[:, :d]
    pt = Y[:, d:2 * d]
    Mt = (Y[:, 2 * d:2 * d + d * d] + 1j * Y[:, 2 * d + d * d:2 * d + 2 * d * d]).reshape(-1, d, d)
    At = Y[:, -2] + 1j * Y[:, -1]
    return xt, pt, Mt, At
