"""Synthetic inputs of the benchmark workloads (host NumPy, seeded).

``packets_2d``: sum of Gaussian-windowed plane-wave packets -- centres uniform
in the inner 80% of the grid, sigma uniform 4-16 cells, |k| uniform
0.1-0.8 Nyquist, direction uniform, amplitude N(0, 1); draws in that order
from ``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

import numpy as np


def packets_2d(N, dx, n_packets=64, seed=0, sigma_cells=(4.0, 16.0), k_frac=(0.1, 0.8)):
    rng = np.random.default_rng(seed)
    n0, n1 = N
    L = np.asarray(N, float) * np.asarray(dx, float)
    c = rng.uniform(0.1, 0.9, (n_packets, 2)) * L
    sig = rng.uniform(sigma_cells[0], sigma_cells[1], n_packets) * min(dx)
    kmag = rng.uniform(k_frac[0], k_frac[1], n_packets) / (2.0 * min(dx))
    th = rng.uniform(0.0, 2 * np.pi, n_packets)
    amp = rng.standard_normal(n_packets)
    x = np.arange(n0)[:, None] * dx[0]
    y = np.arange(n1)[None, :] * dx[1]
    p0 = np.zeros(N)
    for j in range(n_packets):
        rx, ry = x - c[j, 0], y - c[j, 1]
        env = np.exp(-(rx * rx + ry * ry) / (2 * sig[j] ** 2))
        ph = 2 * np.pi * kmag[j] * (np.cos(th[j]) * rx + np.sin(th[j]) * ry)
        p0 += amp[j] * env * np.cos(ph)
    return p0


def phantom_2d(N, dx, seed=0, n_ellipses=(3, 7), edge_cells=1.5, n_packets=2):
    """Ellipses (centres in the inner 60%, semi-axes 4-25% of the grid, random
    orientation, amplitude U(0, 1), logistic edge of width ``edge_cells``) plus
    ``n_packets`` packets of ``packets_2d`` scaled to amplitude 0.3."""
    rng = np.random.default_rng(seed)
    n0, n1 = N
    L = np.asarray(N, float) * np.asarray(dx, float)
    x = np.arange(n0)[:, None] * dx[0]
    y = np.arange(n1)[None, :] * dx[1]
    p0 = np.zeros(N)
    m = int(rng.integers(n_ellipses[0], n_ellipses[1] + 1))
    for _ in range(m):
        c = rng.uniform(0.2, 0.8, 2) * L
        ax = rng.uniform(0.04, 0.25, 2) * L
        th = rng.uniform(0.0, np.pi)
        amp = rng.uniform(0.0, 1.0)
        rx, ry = x - c[0], y - c[1]
        u = (np.cos(th) * rx + np.sin(th) * ry) / ax[0]
        v = (-np.sin(th) * rx + np.cos(th) * ry) / ax[1]
        # signed distance proxy in cells: (1 - rho) * min semi-axis / dx
        rho = np.sqrt(u * u + v * v)
        sd = (1.0 - rho) * min(ax) / min(dx)
        p0 += amp / (1.0 + np.exp(-sd / edge_cells))
    if n_packets:
        p0 += 0.3 * packets_2d(N, dx, n_packets, seed=seed + 7919, sigma_cells=(3.0, 8.0),
                               k_frac=(0.2, 0.6))
    return p0


def packets_3d(N, dx, n_packets=32, seed=0, sigma_cells=(3.0, 10.0), k_frac=(0.1, 0.7)):
    rng = np.random.default_rng(seed)
    L = np.asarray(N, float) * np.asarray(dx, float)
    c = rng.uniform(0.1, 0.9, (n_packets, 3)) * L
    sig = rng.uniform(sigma_cells[0], sigma_cells[1], n_packets) * min(dx)
    kmag = rng.uniform(k_frac[0], k_frac[1], n_packets) / (2.0 * min(dx))
    dirs = rng.standard_normal((n_packets, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    amp = rng.standard_normal(n_packets)
    ax = [np.arange(N[a]) * dx[a] for a in range(3)]
    p0 = np.zeros(N)
    for j in range(n_packets):
        r = [ax[a] - c[j, a] for a in range(3)]
        rx, ry, rz = r[0][:, None, None], r[1][None, :, None], r[2][None, None, :]
        env = np.exp(-(rx * rx + ry * ry + rz * rz) / (2 * sig[j] ** 2))
        ph = 2 * np.pi * kmag[j] * (dirs[j, 0] * rx + dirs[j, 1] * ry + dirs[j, 2] * rz)
        p0 += amp[j] * env * np.cos(ph)
    return p0


_BATCH_LEVELS = {128: (3, (4, 8, 16)), 192: (3, (4, 8, 16)), 256: (3, (4, 8, 16)),
                 384: (3, (4, 8, 16)), 512: (4, (4, 8, 16, 32))}


def batch_forward_problems(device, n_solves=64, seed=0, sizes=(128, 192, 256, 384, 512),
                           nt_range=(400, 2000), beam_range=(1000, 65536)):
    """The batching workload: per solve a square grid drawn from ``sizes``
    (dx 1e-4, c 1500, cfl 0.3; 3 levels (4,8,16), 4 levels (4,8,16,32) at
    512), rho 2 rectangular_mirror MSWPT, top_n drawn log-uniform in
    ``beam_range`` (capped at the half frame), nt uniform in ``nt_range`` on
    the CFL time step, a line sensor on one random edge, and p0 the real part
    of the synthesis of two random unit coefficients (two frames, as the
    128^2 README case), normalised to max |p0| = 1.  Draws from
    ``numpy.random.default_rng(seed)`` in that order, per solve.  MSWPT plans
    are shared between solves of the same grid."""
    import torch
    from .batch import ForwardProblem
    from .decomposition import DyadicDecomposition
    from .geometry import Domain
    from .transforms import MSWPT
    rng = np.random.default_rng(seed)
    plans = {}
    out = []
    for _ in range(n_solves):
        n = int(rng.choice(sizes))
        thr = int(round(float(np.exp(rng.uniform(np.log(beam_range[0]), np.log(beam_range[1]))))))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        edge = int(rng.integers(0, 4))
        if n not in plans:
            lv, nb = _BATCH_LEVELS[n]
            dom = Domain(N=(n, n), dx=(1e-4, 1e-4), c=1500.0, cfl=0.3, periodic=(False, False))
            plans[n] = (dom, MSWPT(DyadicDecomposition(lv, (n, n), nb, (1, 1)), 2,
                                   "rectangular_mirror", device=device))
        dom, wpt = plans[n]
        e = np.zeros(wpt.total_coeffs, dtype=complex)
        e[rng.integers(wpt.coeffs_cumsum[1], wpt.total_coeffs, 2)] = 1.0
        p0 = wpt.inverse(torch.as_tensor(e, device=device)).real.contiguous()
        p0 = p0 / p0.abs().max()
        dt = float(dom.generate_time_domain()[1])
        ts = np.arange(nt) * dt
        j = np.arange(n) * 1e-4
        far = (n - 1) * 1e-4
        line = {0: (np.zeros(n), j), 1: (np.full(n, far), j),
                2: (j, np.zeros(n)), 3: (j, np.full(n, far))}[edge]
        pts = np.stack(line, axis=1)
        out.append(ForwardProblem(p0, dom, wpt, pts, ts, thr))
    return out
