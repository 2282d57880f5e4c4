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
